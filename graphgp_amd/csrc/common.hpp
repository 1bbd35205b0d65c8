// common.hpp -- shared helpers for the gfx950 GraphGP kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>

namespace ggp {

// ------------------------------------------------------------------ errors
// Entry points return 0 on success and a negative status otherwise; the message
// is kept per host thread and read back with ggp_last_error().
enum Status : int {
  kOk = 0,
  kInvalidArgument = -1,
  kHipError = -2,
  kWorkspaceTooSmall = -3,
  kUnsupported = -4,
};

void set_error(const std::string& msg);
const char* last_error();

#define GGP_CHECK_ARG(cond, msg)            \
  do {                                      \
    if (!(cond)) {                          \
      ::ggp::set_error(msg);                \
      return ::ggp::kInvalidArgument;       \
    }                                       \
  } while (0)

#define GGP_HIP(expr)                                                                       \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess) {                                                                 \
      ::ggp::set_error(std::string(#expr) + ": " + hipGetErrorString(e_));                  \
      return ::ggp::kHipError;                                                              \
    }                                                                                       \
  } while (0)

#define GGP_LAUNCH_CHECK()                                                                  \
  do {                                                                                      \
    hipError_t e_ = hipGetLastError();                                                      \
    if (e_ != hipSuccess) {                                                                 \
      ::ggp::set_error(std::string("kernel launch: ") + hipGetErrorString(e_) + " at " +    \
                       __FILE__ + ":" + std::to_string(__LINE__));                         \
      return ::ggp::kHipError;                                                              \
    }                                                                                       \
  } while (0)

#define GGP_TRY(expr)            \
  do {                           \
    int s_ = (expr);             \
    if (s_ != 0) return s_;      \
  } while (0)

// ------------------------------------------------------------------ stage timing
// While stage timing is on for the calling host thread (ggp_stage_profile), a StageScope
// records one event where it is constructed and one where it is destroyed, on stream s,
// so that ggp_stage_times can report the stage's device time; otherwise it does nothing.
// `name` must be a string literal (it is kept by pointer).
class StageScope {
 public:
  StageScope(const char* name, hipStream_t s);
  ~StageScope();
  StageScope(const StageScope&) = delete;
  StageScope& operator=(const StageScope&) = delete;

 private:
  int slot_;
  hipStream_t stream_;
};

// ------------------------------------------------------------------ workspace
// A bump allocator over a caller-provided device buffer (the library never
// allocates).  With base == nullptr it only measures.
struct Workspace {
  char* base;
  size_t cap;
  size_t used = 0;
  Workspace(void* b, size_t c) : base(static_cast<char*>(b)), cap(c) {}
  template <typename T>
  T* take(size_t count) {
    size_t off = (used + 255) & ~size_t(255);
    used = off + count * sizeof(T);
    if (base == nullptr || used > cap) return nullptr;
    return reinterpret_cast<T*>(base + off);
  }
  bool measuring() const { return base == nullptr; }
  bool ok() const { return base != nullptr && used <= cap; }
};

// ------------------------------------------------------------------ k-d tree ids
// The reference tree (graphgp/tree.py:187-208) stores node c's children at
// c + 2^lvl(c) and c + 2^(lvl(c)+1), lvl(c) = floor(log2(c + 1)).
__host__ __device__ inline int tree_level(int64_t c) {
#ifdef __HIP_DEVICE_COMPILE__
  return 63 - __clzll((long long)(c + 1));
#else
  return 63 - __builtin_clzll((unsigned long long)(c + 1));
#endif
}
__host__ __device__ inline int64_t tree_left(int64_t c) { return c + (int64_t(1) << tree_level(c)); }
__host__ __device__ inline int64_t tree_right(int64_t c) { return c + (int64_t(2) << tree_level(c)); }
__host__ __device__ inline int64_t tree_parent(int64_t c) {
  if (c == 0) return -1;
  int lvl = tree_level(c);
  int64_t n_above = (int64_t(1) << lvl) - 1;
  int64_t npl = int64_t(1) << (lvl - 1);
  return c < n_above + npl ? c - npl : c - 2 * npl;
}

// ------------------------------------------------------------------ sort keys
// Order-preserving unsigned images of floats with lax.sort's canonicalisation:
// -0.0 sorts equal to +0.0 and every NaN after +inf.
__host__ __device__ inline uint32_t float_key(float x) {
  uint32_t u;
  if (x == 0.0f) x = 0.0f;
  if (x != x) u = 0x7fc00000u;
  else memcpy(&u, &x, 4);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__host__ __device__ inline uint64_t double_key(double x) {
  uint64_t u;
  if (x == 0.0) x = 0.0;
  if (x != x) u = 0x7ff8000000000000ull;
  else memcpy(&u, &x, 8);
  return (u & 0x8000000000000000ull) ? ~u : (u | 0x8000000000000000ull);
}
__host__ __device__ inline float float_from_key(uint32_t k) {
  uint32_t u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
  float x;
  memcpy(&x, &u, 4);
  return x;
}
__host__ __device__ inline double double_from_key(uint64_t k) {
  uint64_t u = (k & 0x8000000000000000ull) ? (k & 0x7fffffffffffffffull) : ~k;
  double x;
  memcpy(&x, &u, 8);
  return x;
}

template <typename T>
struct KeyOf;
template <>
struct KeyOf<float> {
  using type = uint32_t;
  __host__ __device__ static uint32_t key(float x) { return float_key(x); }
  __host__ __device__ static float value(uint32_t k) { return float_from_key(k); }
};
template <>
struct KeyOf<double> {
  using type = uint64_t;
  __host__ __device__ static uint64_t key(double x) { return double_key(x); }
  __host__ __device__ static double value(uint64_t k) { return double_from_key(k); }
};

// ------------------------------------------------------------------ misc
inline unsigned grid_for(int64_t n, int block) { return (unsigned)((n + block - 1) / block); }
// Grid of a grid-stride launch over n items: a dispatch's size in work-items is a 32-bit
// field, so one-thread-per-item launches must stay below 2^32 threads; 2^20 workgroups is
// plenty to fill the device.
inline unsigned grid_capped(int64_t n, int block) {
  const int64_t g = (n + block - 1) / block;
  return (unsigned)(g < (int64_t(1) << 20) ? g : (int64_t(1) << 20));
}

// Fill a buffer with a 32-bit pattern (used instead of hipMemsetAsync so that a
// captured graph never holds memset nodes).
__global__ void fill_u32_kernel(uint32_t* p, uint32_t v, int64_t n);
int fill_u32(void* p, uint32_t v, int64_t n, hipStream_t s);
int fill_u64(void* p, uint64_t v, int64_t n, hipStream_t s);

// Copy a few bytes (at most 4096) of device memory to the host and wait for them.  The
// copy goes through a pinned buffer of the calling host thread: the runtime stages copies
// into pageable memory through a buffer that all threads share, so such a copy can wait
// on another thread's unrelated work.
int read_to_host(void* dst, const void* src, size_t bytes, hipStream_t s);

inline int dtype_size(int dtype) { return dtype == 1 ? 8 : 4; }  // 0 = float32, 1 = float64

// CU budget of the calling host thread for persistent (work-queue) launches; 0 = all CUs.
int& cu_budget();

// Yield setting of the calling host thread (ggp_set_yield): while the word at `host`
// (pinned host memory, `dev` its device address) is nonzero, the workgroups of persistent
// launches beyond the first keep_cus CUs' worth stop taking work and exit, and grid-stride
// passes launched meanwhile stay within keep_cus CUs.
struct YieldCfg {
  const volatile uint32_t* host = nullptr;
  const uint32_t* dev = nullptr;
  int keep_cus = 0;
};
YieldCfg& yield_cfg();
int device_cus();

// Workgroups of `kernel` (256 threads, lds_bytes of LDS) resident per CU.
inline int blocks_per_cu(const void* kernel, size_t lds_bytes) {
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 256, lds_bytes) != hipSuccess || per_cu < 1)
    per_cu = 1;
  return per_cu;
}

// Grid of a persistent launch: at most `grid` workgroups and, under a CU budget, at most
// budget x (workgroups of `kernel` resident per CU), so the launch leaves room elsewhere.
inline int persistent_grid(const void* kernel, int grid, size_t lds_bytes) {
  const int budget = cu_budget();
  if (budget <= 0) return grid;
  const int cap = budget * blocks_per_cu(kernel, lds_bytes);
  return grid < cap ? grid : cap;
}

// Workgroups of a persistent launch of `grid` (spread over the whole device) that never
// yield: keep_cus CUs' share of them, at least one.
inline uint32_t yield_keep(int grid) {
  const YieldCfg& y = yield_cfg();
  if (y.dev == nullptr) return (uint32_t)grid;
  const int cus = device_cus();
  const int64_t keep = ((int64_t)grid * y.keep_cus + cus - 1) / cus;
  return (uint32_t)(keep < 1 ? 1 : (keep > grid ? grid : keep));
}

// Whether workgroup blockIdx.x of a persistent launch should stop taking work now.
__device__ inline bool yield_now(const uint32_t* flag, uint32_t keep) {
  return flag != nullptr && blockIdx.x >= keep &&
         __hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) != 0u;
}

// Grid of a grid-stride launch, capped under the calling thread's CU budget like a
// persistent launch, and to the yield share while the yield word is set (long grid-stride
// kernels of a large job then leave room elsewhere).
inline unsigned budget_grid(const void* kernel, unsigned grid) {
  int g = persistent_grid(kernel, (int)(grid < (1u << 30) ? grid : (1u << 30)), 0);
  const YieldCfg& y = yield_cfg();
  if (y.host != nullptr && *y.host != 0u) {
    const int cap = y.keep_cus * blocks_per_cu(kernel, 0);
    if (cap >= 1 && g > cap) g = cap;
  }
  return (unsigned)g;
}

}  // namespace ggp
