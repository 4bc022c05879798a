// fwx_profile.hip -- optional per-kernel-class timing with HIP events.
//
// While enabled (fwx_profile_begin .. fwx_profile_end), the drivers bracket
// each launch group of a class with a pair of events recorded on the stream
// the kernels are launched on, and note the group's algorithmic work; the
// pairs are resolved at fwx_profile_end.  Solves launch directly (no graph
// replay) while profiling is on, so every event is a live stream event.
// Off by default: a disabled profiler costs one relaxed atomic load per
// launch group.
#include <hip/hip_runtime.h>

#include <atomic>
#include <mutex>
#include <vector>

#include "fwx_internal.hpp"

namespace fwx {

namespace {

struct ProfRec {
  int cls;
  hipEvent_t a, b;
  double work;
};

std::atomic<bool> g_on{false};
std::mutex g_mu;
std::vector<ProfRec> g_recs;
double g_ms[FWX_PROF_NCLASS];
double g_work[FWX_PROF_NCLASS];
int64_t g_groups[FWX_PROF_NCLASS];

void clear_totals() {
  for (int c = 0; c < FWX_PROF_NCLASS; ++c) {
    g_ms[c] = 0.0;
    g_work[c] = 0.0;
    g_groups[c] = 0;
  }
}

}  // namespace

bool profiling() { return g_on.load(std::memory_order_relaxed); }

int prof_open(hipStream_t st, hipEvent_t* a) {
  *a = nullptr;
  if (!profiling()) return FWX_OK;
  FWX_HIP(hipEventCreate(a));
  FWX_HIP(hipEventRecord(*a, st));
  return FWX_OK;
}

int prof_close(hipStream_t st, hipEvent_t a, int cls, double work) {
  if (!a) return FWX_OK;
  hipEvent_t b;
  FWX_HIP(hipEventCreate(&b));
  FWX_HIP(hipEventRecord(b, st));
  std::lock_guard<std::mutex> lk(g_mu);
  g_recs.push_back(ProfRec{cls, a, b, work});
  return FWX_OK;
}

static void prof_release(std::vector<ProfRec>& recs) {
  for (ProfRec& r : recs) {
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  recs.clear();
}

}  // namespace fwx

using namespace fwx;

extern "C" {

int fwx_profile_begin(void) {
  clear_error();
  std::lock_guard<std::mutex> lk(g_mu);
  prof_release(g_recs);
  clear_totals();
  g_on.store(true);
  return FWX_OK;
}

int fwx_profile_end(double* ms, double* work, int64_t* groups) {
  clear_error();
  g_on.store(false);
  std::lock_guard<std::mutex> lk(g_mu);
  int rc = FWX_OK;
  for (ProfRec& r : g_recs) {
    float t = 0.f;
    if (rc == FWX_OK) {
      hipError_t e = hipEventSynchronize(r.b);
      if (e == hipSuccess) e = hipEventElapsedTime(&t, r.a, r.b);
      if (e != hipSuccess) rc = hip_fail(e, "hipEventElapsedTime", __FILE__, __LINE__);
    }
    if (r.cls >= 0 && r.cls < FWX_PROF_NCLASS) {
      g_ms[r.cls] += t;
      g_work[r.cls] += r.work;
      g_groups[r.cls] += 1;
    }
  }
  prof_release(g_recs);
  for (int c = 0; c < FWX_PROF_NCLASS; ++c) {
    if (ms) ms[c] = g_ms[c];
    if (work) work[c] = g_work[c];
    if (groups) groups[c] = g_groups[c];
  }
  clear_totals();
  return rc;
}

}  // extern "C"
