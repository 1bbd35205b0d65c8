// refine_launch.hpp -- instantiation helper for the refine kernel configurations.
#pragma once
#include "refine_impl.hpp"

namespace ggp {

size_t refine_lds_limit();

template <typename T, int KK, int P, int D, bool FAST, int MODE>
int launch_variant(const RefineArgs<T>& a, hipStream_t s, size_t lds_bytes, int grid) {
  auto kern = refine_kernel<T, KK, P, D, FAST, MODE>;
  if (lds_bytes > 65536) {
    hipError_t e = hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
    if (e != hipSuccess) {
      set_error(std::string("refine: cannot reserve LDS for the covariance table: ") + hipGetErrorString(e));
      return kUnsupported;
    }
  }
  grid = persistent_grid((const void*)kern, grid, lds_bytes);
  if (a.reserve_wgs > 0) {
    const int cap = blocks_per_cu((const void*)kern, lds_bytes) * device_cus() - a.reserve_wgs;
    if (cap >= 1 && grid > cap) grid = cap;
  }
  if (a.grid_limit > 0 && grid > a.grid_limit) grid = a.grid_limit;
  RefineArgs<T> b = a;
  b.yield = yield_cfg().dev;
  b.keep = yield_keep(grid);
  hipLaunchKernelGGL(kern, dim3(grid), dim3(256), lds_bytes, s, b);
  GGP_LAUNCH_CHECK();
  return kOk;
}

// Launches the fast-lookup variant (log-spaced bins) and the exact variant; the table
// check written by the prep kernel makes exactly one of them do the work.
template <typename T, int KK, int P, int D>
int launch_config(RefineArgs<T> a, hipStream_t s, size_t lds_bytes, int grid) {
  constexpr int64_t gpt = 256 / P;
  a.ntiles = ((a.list_in ? (int64_t)a.fix_cap : a.n_ref) + gpt - 1) / gpt;
  if (a.ntiles == 0) return kOk;
  if ((int64_t)grid > a.ntiles) grid = (int)a.ntiles;
  // the FP64 recomputation needs one KK x (KK+1) FP64 slab per recomputing wavefront and
  // KK FP64 weights per point group after the table
  // (the weights mode only lists the points; it needs no slabs)
  if (a.tab64 && a.mode != kWeights)
    lds_bytes = (size_t)a.fix_lds_offset + (size_t)kFixWaves * KK * (KK + 1) * sizeof(double) +
                (size_t)(256 / P) * KK * sizeof(double);
  if (lds_bytes > refine_lds_limit()) {
    set_error("refine: covariance table does not fit in LDS (" + std::to_string(lds_bytes) + " bytes)");
    return kUnsupported;
  }
  // one instantiation per mode, without the other modes' code (fewer live values, fewer
  // spills than one kernel that reads the mode at run time)
  switch (a.mode) {
    case kWeights:
      GGP_TRY((launch_variant<T, KK, P, D, true, kWeights>(a, s, lds_bytes, grid)));
      return launch_variant<T, KK, P, D, false, kWeights>(a, s, lds_bytes, grid);
    case kGenerate:
      GGP_TRY((launch_variant<T, KK, P, D, true, kGenerate>(a, s, lds_bytes, grid)));
      return launch_variant<T, KK, P, D, false, kGenerate>(a, s, lds_bytes, grid);
    case kInverse:
      GGP_TRY((launch_variant<T, KK, P, D, true, kInverse>(a, s, lds_bytes, grid)));
      return launch_variant<T, KK, P, D, false, kInverse>(a, s, lds_bytes, grid);
    case kLogdet:
      GGP_TRY((launch_variant<T, KK, P, D, true, kLogdet>(a, s, lds_bytes, grid)));
      return launch_variant<T, KK, P, D, false, kLogdet>(a, s, lds_bytes, grid);
    default:
      set_error("refine: unknown mode " + std::to_string(a.mode));
      return kInvalidArgument;
  }
}


}  // namespace ggp
