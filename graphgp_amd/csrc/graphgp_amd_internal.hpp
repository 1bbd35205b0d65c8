// graphgp_amd_internal.hpp -- internal entry points behind the C ABI (abi.cpp).
#pragma once
#include "common.hpp"

namespace ggp {

struct TabMeta;

size_t build_tree_workspace(int64_t n, int d, int dtype);
int build_tree(const void* pts, int64_t n, int d, int dtype, void* pts_out, int32_t* split_dims, int32_t* indices,
               void* ws, size_t ws_bytes, hipStream_t s);

size_t query_neighbors_workspace(int64_t n, int d, int64_t nq, int k, int dtype);
size_t graph_schedule_workspace(int64_t n_ref);
int graph_schedule(const void* pts, int dtype, int64_t n, int d, const int32_t* nbrs, int64_t n0, int k,
                   const int64_t* offsets, int64_t n_off, int32_t* sched, int32_t* nbv, void* pts_sched, void* ws,
                   size_t ws_bytes, hipStream_t s);
int query_neighbors(const void* pts, const int32_t* split, int64_t n, int d, int dtype, const int32_t* qidx,
                    const int32_t* midx, int64_t q0, int64_t nq, int k, int32_t* out, void* ws, size_t ws_bytes,
                    hipStream_t s);

size_t compute_depths_workspace(int64_t n_ref, int k);
enum DepthsMethod : int { kDepthsFlow = 0, kDepthsRounds = 1, kDepthsSweeps = 2 };
int compute_depths(const int32_t* nbrs, int64_t n_ref, int k, int64_t n0, int32_t* depths, int method, void* ws,
                   size_t ws_bytes, hipStream_t s);
size_t order_by_depth_workspace(int64_t n);
int order_by_depth(const void* pts, int64_t n, int d, int dtype, const int32_t* idx, const int32_t* nbrs,
                   int64_t n_ref, int k, const int32_t* depths, int32_t max_depth, void* pts_o, int32_t* idx_o,
                   int32_t* nbrs_o, int32_t* dep_o, void* ws, size_t ws_bytes, hipStream_t s);
int depth_offsets(const int32_t* depths_sorted, int64_t n, int64_t n_off, int64_t* offsets, hipStream_t s);
int check_graph(const int32_t* nbrs, int64_t n_ref, int k, int64_t n, const int64_t* offsets, int64_t n_off,
                uint32_t* err_dev, hipStream_t s);

size_t table_workspace(int nb, int t64);
int prep_table(const double* bins, const double* vals, int nb, int t64, int increment, void* tab, TabMeta* meta,
               hipStream_t s);

}  // namespace ggp
