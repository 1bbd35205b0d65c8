// dense.hpp -- FP64 seed-block routines (dense.hip).
#pragma once
#include "common.hpp"

namespace ggp {
size_t dense_workspace(int64_t n0);
// A <- chol(K(points[0:n0])) (lower, row-major n0 x n0).
int dense_factor(const void* pts, int pts_f64, int d, int64_t n0, const double* bins, const double* vals, int nb,
                 double* A, hipStream_t s);
int dense_apply(const double* L, int64_t n0, const double* xs, double* out, hipStream_t s);  // out = L xs
int dense_solve(const double* L, int64_t n0, double* v, hipStream_t s);                     // v <- L^{-1} v
int cov_matrix(const void* a, int64_t na, const void* b, int64_t nb, int d, int f64, const double* bins,
               const double* vals, int nbins, double* out, hipStream_t s);
int dense_logdiag(const double* L, int64_t n0, double* out, hipStream_t s);                 // sum log L_ii
// When *logdet (from dense_logdiag) is not finite -- the Cholesky failed -- rebuild K in A
// and overwrite *logdet with log|det K| / 2 from a partially pivoted LU (slogdet's value);
// otherwise both kernels return at once.
int dense_logdet_fallback(const void* pts, int pts_f64, int d, int64_t n0, const double* bins, const double* vals,
                          int nb, double* A, double* logdet, hipStream_t s);
}  // namespace ggp
