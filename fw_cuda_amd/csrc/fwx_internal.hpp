// fwx_internal.hpp -- host-side plumbing shared by the libfwx translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/fwx.h"

namespace fwx {

// Thread-local last-error message (fwx_last_error).
void set_error(const std::string& msg);
void clear_error();

int fail(int code, const std::string& msg);
int hip_fail(hipError_t e, const char* what, const char* file, int line);

#define FWX_HIP(call)                                                  \
  do {                                                                 \
    hipError_t fwx_e_ = (call);                                        \
    if (fwx_e_ != hipSuccess) return ::fwx::hip_fail(fwx_e_, #call, __FILE__, __LINE__); \
  } while (0)

#define FWX_LAUNCHED(what)                                             \
  do {                                                                 \
    hipError_t fwx_e_ = hipGetLastError();                             \
    if (fwx_e_ != hipSuccess) return ::fwx::hip_fail(fwx_e_, what, __FILE__, __LINE__); \
  } while (0)

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Plain-FW drivers (enqueue only; the C entry points add the final sync).
int enqueue_baseline(int32_t* d, int32_t n, hipStream_t st);
int enqueue_blocked(int32_t* d, int32_t n, int32_t tile, int32_t kappa, hipStream_t st);
// A plain solve of a matrix whose cells are all in [0, 2^30), with the tuned
// configuration and no pk16 passes (the exact finish of a batched graph).
int enqueue_blocked_exact(int32_t* d, int32_t n, hipStream_t st);
// The form of this thread's last enqueue_blocked (fwx_last_solve_form).
void last_solve_form(int32_t* form, int32_t* exact_finish);
// 1 when that solve's pk16 rest passes ran on packed stage panels.
int last_solve_packed();
int enqueue_refshape(int32_t* d, int32_t n, hipStream_t st);
// The fused rounds of a small solve on 32-wide k-tiles (fwx_small.hip), and
// the region width they use for size n (a graph-cache key).
int enqueue_small_rounds(int32_t* d, int32_t n, hipStream_t st);
int small_region_knob(int32_t n);

// Event profiling (fwx_profile.hip): prof_open records a start event when
// profiling is on (else leaves *a null); prof_close records the end event.
// Profiled work must be launched directly, never captured into a graph.
bool profiling();
// Flags for every event the library creates (timing or ordering only).
unsigned event_flags(bool timing);
int prof_open(hipStream_t st, hipEvent_t* a);
int prof_close(hipStream_t st, hipEvent_t a, int cls, double work);
// prof_chain closes the group started at *a (if profiling) with a new event
// and makes that event the start of the next group on the same stream: one
// marker between two adjacent groups instead of two.
int prof_chain(hipStream_t st, hipEvent_t* a, int cls, double work);
// Records a fresh timing event on st (released at fwx_profile_end).
int prof_mark(hipStream_t st, hipEvent_t* e);

// A device int from a file-scope array (fwx_range.hip), held by one call:
// flag_acquire takes a free slot of the current device and zeroes it on st;
// kernels OR verdicts into *f; flag_release reads it into *value
// (synchronising st) and frees the slot.  With value null it only waits for
// st to drain and frees the slot (the path after a failed launch).
struct DeviceFlag {
  int* f = nullptr;
  int dev = -1;
  int slot = -1;
};
int flag_acquire(hipStream_t st, DeviceFlag* fl);
int flag_release(hipStream_t st, DeviceFlag* fl, int* value);

// Range check of count cells at d (fwx_range.hip): *ok = every cell in
// [0, 2^30), i.e. the pk64 relaxation form is exact for the whole run.
// Synchronises `st`.
int values_in_pk64_range(const int32_t* d, int64_t count, hipStream_t st, bool* ok);
// *ok = every one of count cells at d is below cap (fwx_range.hip); the
// check after a solve with pk16 rest passes.  Synchronises `st`.
int values_below(const int32_t* d, int64_t count, int32_t cap, hipStream_t st, bool* ok);
// The two checks of a solve that does not wait for its range check
// (fwx_range.hip).  range_check_enqueue zeroes flags[0] and flags[1] on st
// and ORs "a cell outside [0, 2^30)" into flags[0], without synchronising.
// deferred_checks_read, after the solve, ORs "a cell >= cap" into flags[1]
// by a pass over the matrix if `scan` (else the solve's kernels have done
// it), copies both words into out and synchronises st.  `flags` is two
// device ints that the caller holds for the whole solve.
int range_check_enqueue(const int32_t* d, int64_t count, int* flags, hipStream_t st);
int deferred_checks_read(const int32_t* d, int64_t count, int32_t cap, bool scan, int* flags, hipStream_t st,
                         int out[2]);
// The per-device second stream of the look-ahead schedules (fwx_blocked.hip).
int side_stream(hipStream_t* out);
// Any D[i][i] < 0 on a device matrix (fwx_range.hip); synchronises `st`.
int device_negative_cycle(const int32_t* d, int32_t n, hipStream_t st, bool* neg);

// Tuned (tile, kappa) for a plain run of size n, and the shape of its
// 128-wide rest kernel (0, 1 or 2, see fwx_blocked.hip; pk64 and min3 runs).
void tuned_config(int32_t n, int32_t* tile, int32_t* kappa);
int tuned_rest_variant(int32_t n);
// Chunk depth (16 or 32) of a pk16 128-wide rest pass of size n.
int tuned_p16_chunk(int32_t n);
// Tuned tile of tiling_gpu (kappa 1) for size n.
int tuned_tiling_tile(int32_t n);
// Tuned two-stream schedule (FWX_LOOKAHEAD values 1..3) for size n.
int tuned_lookahead(int32_t n);
// Tuned (tile, kappa) of a graph of size n inside a batch, and the threads of
// its 128-wide products (512: 8x4 cells each; 256: 8x8).
void batch_tuned_config(int32_t n, int32_t* tile, int32_t* kappa, int32_t* rest_threads);

}  // namespace fwx
