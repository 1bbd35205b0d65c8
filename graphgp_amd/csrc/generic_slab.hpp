// generic_slab.hpp -- dense per-point linear algebra for the one-lane-per-point kernels
// (any k, any d), on a system held in device memory.
//
// Each lane owns one point's (k+1) x (k+1) system.  The 64 lanes of a block store their
// systems interleaved (element e of lane l at e * 64 + l), so the lanes' accesses to the
// same element coalesce into one request.  The factorisation is the dot-product
// (Cholesky-Crout) form: every entry is one dot product of two rows, whose loads are
// independent of each other, so four partial sums keep several loads in flight (the
// right-looking form chains a load, an FMA and a store per step and waits on memory at
// every one).
#pragma once

#include "common.hpp"

namespace ggp {

template <typename T>
struct LaneSlab {
  T* base;  // this lane's element 0
  __device__ __forceinline__ T& operator[](size_t e) const { return base[e * 64]; }
};

// sum_{c < n} x[xo + c] * y[yo + c] with four independent partial sums
template <typename T>
__device__ __forceinline__ T slab_dot(const LaneSlab<T>& s, size_t xo, size_t yo, int n) {
  T s0 = T(0), s1 = T(0), s2 = T(0), s3 = T(0);
  int c = 0;
  for (; c + 4 <= n; c += 4) {
    const T a0 = s[xo + c], a1 = s[xo + c + 1], a2 = s[xo + c + 2], a3 = s[xo + c + 3];
    const T b0 = s[yo + c], b1 = s[yo + c + 1], b2 = s[yo + c + 2], b3 = s[yo + c + 3];
    s0 += a0 * b0;
    s1 += a1 * b1;
    s2 += a2 * b2;
    s3 += a3 * b3;
  }
  for (; c < n; ++c) s0 += s[xo + c] * s[yo + c];
  return (s0 + s1) + (s2 + s3);
}

// Lower Cholesky, in place, of the leading n x n block of the row-major matrix at offset 0
// with row stride ld (only the lower triangle is read or written).
template <typename T>
__device__ __forceinline__ void slab_cholesky(const LaneSlab<T>& s, int n, int ld) {
  for (int i = 0; i < n; ++i) {
    const size_t ri = (size_t)i * ld;
    for (int j = 0; j < i; ++j) {
      const size_t rj = (size_t)j * ld;
      s[ri + j] = (s[ri + j] - slab_dot(s, ri, rj, j)) / s[rj + j];
    }
    s[ri + i] = sqrt(s[ri + i] - slab_dot(s, ri, ri, i));
  }
}

// x <- L^{-1} x (x at offset xo, n entries; L the factor above)
template <typename T>
__device__ __forceinline__ void slab_solve_lower(const LaneSlab<T>& s, int n, int ld, size_t xo) {
  for (int c = 0; c < n; ++c) {
    const size_t rc = (size_t)c * ld;
    s[xo + c] = (s[xo + c] - slab_dot(s, rc, xo, c)) / s[rc + c];
  }
}

// x <- L^{-T} x
template <typename T>
__device__ __forceinline__ void slab_solve_upper(const LaneSlab<T>& s, int n, int ld, size_t xo) {
  for (int c = n - 1; c >= 0; --c) {
    T a0 = T(0), a1 = T(0);
    int j = c + 1;
    for (; j + 2 <= n; j += 2) {
      a0 += s[(size_t)j * ld + c] * s[xo + j];
      a1 += s[(size_t)(j + 1) * ld + c] * s[xo + j + 1];
    }
    if (j < n) a0 += s[(size_t)j * ld + c] * s[xo + j];
    s[xo + c] = (s[xo + c] - (a0 + a1)) / s[(size_t)c * ld + c];
  }
}

}  // namespace ggp
