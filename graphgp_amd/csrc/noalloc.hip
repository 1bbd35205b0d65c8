// noalloc.hip -- generate_logdet with no workspace argument and no allocation.
//
// For callers that cannot allocate device memory (code under HIP graph capture, tight
// serving or fitting loops, harnesses that forbid hipMalloc): the scratch of the log-det
// flow (counters, prepared covariance table, per-tile partial sums, the FP64 seed block)
// lives in one file-scope device array of kNoallocScratchBytes per GPU, held until the
// process ends.  The kernels are those of ggp_generate_logdet; only the scratch differs.
//
// The array's name is a host-side stand-in, not a device address: its address on the
// current device comes from hipGetSymbolAddress, looked up once per device and kept.
// Every call on a device shares that device's array, so calls that overlap on one GPU
// (two streams, two host threads) race; the flow clears its counters itself at the start
// of every call, and the partial sums are overwritten tile by tile.
#include <mutex>
#include <string>

#include "../../include/graphgp_amd.h"
#include "common.hpp"
#include "refine_ops.hpp"

namespace ggp {
namespace {

constexpr size_t kNoallocScratchBytes = (size_t)GGP_NOALLOC_SCRATCH_BYTES;
static_assert(kNoallocScratchBytes % 256 == 0, "scratch size must keep 256-byte alignment");

__device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device (looked up on the first call there).
int scratch_address(void** out) {
  int dev = -1;
  GGP_HIP(hipGetDevice(&dev));
  GGP_CHECK_ARG(dev >= 0 && dev < kMaxDevices, "generate_logdet_noalloc: device ordinal out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr) {
      set_error(std::string("generate_logdet_noalloc: hipGetSymbolAddress(scratch) failed: ") +
                hipGetErrorString(e));
      return kHipError;
    }
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return kOk;
}

}  // namespace
}  // namespace ggp

extern "C" int ggp_generate_logdet_noalloc(const void* points, int pdtype, int64_t n, int d,
                                           const int32_t* neighbors, int64_t n0, int k, const double* cov_bins,
                                           const double* cov_vals, int n_bins, double* logdet_device, int precision,
                                           void* stream) {
  using namespace ggp;
  GGP_CHECK_ARG(n >= 1 && n0 >= 1 && n0 <= n && k >= 1 && d >= 1 && n_bins >= 2,
                "generate_logdet_noalloc: bad sizes");
  RefineProblem p{};
  p.n = n;
  p.d = d;
  p.n_ref = n - n0;
  p.k = k;
  p.nb = n_bins;
  p.increment = precision == GGP_PREC_FP32_INC;
  p.with_seed = 1;
  const size_t need = refine_workspace(p, 2, precision == GGP_PREC_FP64 ? 1 : 0);
  if (need > kNoallocScratchBytes) {
    set_error("generate_logdet_noalloc: this problem needs " + std::to_string(need) +
              " bytes of scratch, more than the " + std::to_string(kNoallocScratchBytes) +
              " bytes of the static buffer; use ggp_generate_logdet with a workspace");
    return kInvalidArgument;
  }
  void* ws = nullptr;
  GGP_TRY(scratch_address(&ws));
  return ggp_generate_logdet(points, pdtype, n, d, neighbors, n0, k, cov_bins, cov_vals, n_bins, logdet_device,
                             precision, ws, kNoallocScratchBytes, stream);
}
