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
#include <cstdlib>
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
std::vector<hipEvent_t> g_events;  // every event created while profiling
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

// Every event here orders or times work on one device, so none needs the
// default system-scope release (a full L2 writeback and invalidate, about
// 10 us per record on a solve that keeps a 256 MiB matrix hot).  Events are
// created with hipEventDisableSystemFence: the device-scope fence of an
// ordinary kernel boundary.  FWX_EVENT_SYSFENCE=1 restores the default.
unsigned event_flags(bool timing) {
  const char* e = std::getenv("FWX_EVENT_SYSFENCE");
  const bool sys = e && e[0] == '1';
  return (timing ? 0u : (unsigned)hipEventDisableTiming) | (sys ? 0u : (unsigned)hipEventDisableSystemFence);
}

int prof_open(hipStream_t st, hipEvent_t* a) {
  *a = nullptr;
  if (!profiling()) return FWX_OK;
  return prof_mark(st, a);
}

int prof_mark(hipStream_t st, hipEvent_t* e) {
  FWX_HIP(hipEventCreateWithFlags(e, event_flags(true)));
  {
    std::lock_guard<std::mutex> lk(g_mu);
    g_events.push_back(*e);
  }
  FWX_HIP(hipEventRecord(*e, st));
  return FWX_OK;
}

int prof_close(hipStream_t st, hipEvent_t a, int cls, double work) {
  if (!a) return FWX_OK;
  hipEvent_t b;
  if (int rc = prof_mark(st, &b)) return rc;
  std::lock_guard<std::mutex> lk(g_mu);
  g_recs.push_back(ProfRec{cls, a, b, work});
  return FWX_OK;
}

int prof_chain(hipStream_t st, hipEvent_t* a, int cls, double work) {
  if (!*a) return FWX_OK;
  hipEvent_t b;
  if (int rc = prof_mark(st, &b)) return rc;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    g_recs.push_back(ProfRec{cls, *a, b, work});
  }
  *a = b;
  return FWX_OK;
}

// Caller holds g_mu.
static void prof_release() {
  for (hipEvent_t e : g_events) (void)hipEventDestroy(e);
  g_events.clear();
  g_recs.clear();
}

}  // namespace fwx

using namespace fwx;

extern "C" {

int fwx_profile_begin(void) {
  clear_error();
  std::lock_guard<std::mutex> lk(g_mu);
  prof_release();
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
  prof_release();
  for (int c = 0; c < FWX_PROF_NCLASS; ++c) {
    if (ms) ms[c] = g_ms[c];
    if (work) work[c] = g_work[c];
    if (groups) groups[c] = g_groups[c];
  }
  clear_totals();
  return rc;
}

}  // extern "C"
