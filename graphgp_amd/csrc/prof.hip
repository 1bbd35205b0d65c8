// prof.hip -- stage timing: while switched on for a host thread, each StageScope records a
// pair of hipEvents around one stage of an entry point, on the stream the stage runs on;
// ggp_stage_times reads them back (waiting on the events) and sums them per stage name.
#include <vector>

#include "common.hpp"
#include "graphgp_amd_internal.hpp"

namespace ggp {

namespace {

struct StageRec {
  const char* name;
  hipEvent_t start, stop;
};
struct StageSum {
  const char* name;
  double ms;
  int count;
};

thread_local bool t_stage_on = false;
thread_local std::vector<StageRec> t_pending;
thread_local std::vector<hipEvent_t> t_pool;  // events of read-back records, for reuse
thread_local std::vector<StageSum> t_sums;

hipEvent_t take_event() {
  if (!t_pool.empty()) {
    hipEvent_t e = t_pool.back();
    t_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  return hipEventCreate(&e) == hipSuccess ? e : nullptr;
}

// Read every pending record into the sums (in order of first appearance of each name).
int collect() {
  int rc = kOk;
  for (const StageRec& r : t_pending) {
    float ms = 0.f;
    if (rc == kOk) {
      hipError_t e = hipEventSynchronize(r.stop);
      if (e == hipSuccess) e = hipEventElapsedTime(&ms, r.start, r.stop);
      if (e != hipSuccess) {
        set_error(std::string("stage timing: ") + hipGetErrorString(e));
        rc = kHipError;
      }
    }
    if (rc == kOk) {
      StageSum* hit = nullptr;
      for (StageSum& s : t_sums)
        if (std::strcmp(s.name, r.name) == 0) hit = &s;
      if (hit == nullptr) {
        t_sums.push_back({r.name, 0.0, 0});
        hit = &t_sums.back();
      }
      hit->ms += ms;
      hit->count += 1;
    }
    t_pool.push_back(r.start);
    t_pool.push_back(r.stop);
  }
  t_pending.clear();
  return rc;
}

}  // namespace

StageScope::StageScope(const char* name, hipStream_t s) : slot_(-1), stream_(s) {
  if (!t_stage_on) return;
  hipEvent_t a = take_event(), b = take_event();
  if (a == nullptr || b == nullptr || hipEventRecord(a, s) != hipSuccess) {
    if (a) t_pool.push_back(a);
    if (b) t_pool.push_back(b);
    return;  // timing only: the stage itself runs regardless
  }
  slot_ = (int)t_pending.size();
  t_pending.push_back({name, a, b});
}

StageScope::~StageScope() {
  if (slot_ >= 0) (void)hipEventRecord(t_pending[slot_].stop, stream_);
}

}  // namespace ggp

extern "C" {

int ggp_stage_profile(int enable) {
  using namespace ggp;
  const int rc = collect();  // pending records of an earlier period are dropped below
  t_sums.clear();
  t_stage_on = enable != 0;
  return rc;
}

int ggp_stage_times(int capacity, const char** names, double* total_ms, int* counts) {
  using namespace ggp;
  GGP_TRY(collect());
  const int n = (int)t_sums.size();
  for (int i = 0; i < n && i < capacity; ++i) {
    if (names) names[i] = t_sums[i].name;
    if (total_ms) total_ms[i] = t_sums[i].ms;
    if (counts) counts[i] = t_sums[i].count;
  }
  return n;
}

}  // extern "C"
