// fwx_range.hip -- per-call choice of the relaxation form.
//
// The pk64 form (fwx_device.hpp) is exact when every operand lies in
// [0, 2^30).  Floyd-Warshall only ever stores min(old, a + b) where a, b are
// earlier values, so if every input cell is in that range every value of the
// run is too.  One pass over the input decides the form for the whole call.
//
// The flag lives in a file-scope device array (no allocation).  A call holds
// one of its 64 slots from the zero fill until its result is back on the
// host; a 65th concurrent call waits for a slot to be released, so no two
// calls in flight ever share one.
//
// A solve on packed stage panels does not wait for the pass: it enqueues it
// into a word of its own workspace (range_check_enqueue), every kernel of
// that solve returns at entry when the word is set, and one read-back after
// the solve brings the word back with the cap check's (deferred_checks_read).
#include <hip/hip_runtime.h>

#include <condition_variable>
#include <mutex>

#include "fwx_device.hpp"
#include "fwx_internal.hpp"

namespace fwx {

namespace {

constexpr int kSlots = 64;
__device__ int g_range_flags[kSlots];

__global__ void k_flag_zero(int* f) { *f = 0; }

__global__ __launch_bounds__(256) void k_range(const int32_t* __restrict__ d, int64_t count, int vec4,
                                               int* __restrict__ f) {
  uint32_t o = 0;
  const int64_t tid = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * 256;
  int64_t done = 0;
  if (vec4) {
    const int4* d4 = reinterpret_cast<const int4*>(d);
    const int64_t n4 = count / 4;
    for (int64_t i = tid; i < n4; i += stride) {
      const int4 v = d4[i];
      o |= (uint32_t)v.x | (uint32_t)v.y | (uint32_t)v.z | (uint32_t)v.w;
    }
    done = n4 * 4;
  }
  for (int64_t i = done + tid; i < count; i += stride) o |= (uint32_t)d[i];
  const bool bad = (o & kPk64Mask) != 0u;
  if (__ballot(bad) != 0ull && (threadIdx.x & 63) == 0) atomicOr(f, 1);
}

std::mutex g_mu;
int* g_flag_base[64] = {};
// Slots in use, per device (bit s = slot s), and the wait for a free one.
uint64_t g_busy[64] = {};
std::condition_variable g_freed;

int flag_base(int** out) {
  int dev = 0;
  FWX_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) return fail(FWX_ERR_UNSUPPORTED, "device index above 63");
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_flag_base[dev]) {
    void* p = nullptr;
    FWX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(g_range_flags)));
    g_flag_base[dev] = static_cast<int*>(p);
  }
  *out = g_flag_base[dev];
  return FWX_OK;
}

// A slot of the current device's flag array, held until release_slot.
int acquire_slot(int** f, int* dev_out, int* slot_out) {
  int* base = nullptr;
  if (int rc = flag_base(&base)) return rc;
  int dev = 0;
  FWX_HIP(hipGetDevice(&dev));
  std::unique_lock<std::mutex> lk(g_mu);
  g_freed.wait(lk, [&] { return g_busy[dev] != ~0ull; });
  const int s = __builtin_ctzll(~g_busy[dev]);
  g_busy[dev] |= 1ull << s;
  *f = base + s;
  *dev_out = dev;
  *slot_out = s;
  return FWX_OK;
}

void release_slot(int dev, int s) {
  {
    std::lock_guard<std::mutex> lk(g_mu);
    g_busy[dev] &= ~(1ull << s);
  }
  g_freed.notify_one();
}

}  // namespace

int flag_acquire(hipStream_t st, DeviceFlag* fl) {
  if (int rc = acquire_slot(&fl->f, &fl->dev, &fl->slot)) return rc;
  k_flag_zero<<<1, 1, 0, st>>>(fl->f);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    release_slot(fl->dev, fl->slot);
    fl->f = nullptr;
    return hip_fail(e, "k_flag_zero", __FILE__, __LINE__);
  }
  return FWX_OK;
}

int flag_release(hipStream_t st, DeviceFlag* fl, int* value) {
  if (!fl->f) return FWX_OK;
  int rc = FWX_OK;
  if (value) {
    hipError_t e = hipMemcpyAsync(value, fl->f, sizeof(int), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) rc = hip_fail(e, "flag read-back", __FILE__, __LINE__);
  } else {
    // a launch failed after the zero fill was enqueued: let the stream
    // drain before the slot can be handed out again
    (void)hipStreamSynchronize(st);
  }
  release_slot(fl->dev, fl->slot);
  fl->f = nullptr;
  return rc;
}

namespace {

// Zero a slot, run `launch` (which ORs into it), read it back; the slot is
// released on every path.
template <class F>
int flag_query(hipStream_t st, F launch, int* h) {
  DeviceFlag fl;
  if (int rc = flag_acquire(st, &fl)) return rc;
  const int rc = launch(fl.f);
  const int rc2 = flag_release(st, &fl, rc ? nullptr : h);
  return rc ? rc : rc2;
}

}  // namespace

int values_in_pk64_range(const int32_t* d, int64_t count, hipStream_t st, bool* ok) {
  const int64_t work = (count + 3) / 4;
  const int blocks = (int)std::min<int64_t>((work + 255) / 256, 4096);
  const int vec4 = (reinterpret_cast<uintptr_t>(d) & 15u) == 0;
  int h = 1;
  const int rc = flag_query(
      st,
      [&](int* f) -> int {
        k_range<<<blocks > 0 ? blocks : 1, 256, 0, st>>>(d, count, vec4, f);
        FWX_LAUNCHED("k_range");
        return FWX_OK;
      },
      &h);
  if (rc) return rc;
  *ok = (h == 0);
  return FWX_OK;
}

namespace {

// Any cell >= cap -> *f = 1 (signed compare).
__global__ __launch_bounds__(256) void k_at_least(const int32_t* __restrict__ d, int64_t count, int vec4,
                                                  int32_t cap, int* __restrict__ f) {
  int32_t mx = INT32_MIN;
  const int64_t tid = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * 256;
  int64_t done = 0;
  if (vec4) {
    const int4* d4 = reinterpret_cast<const int4*>(d);
    const int64_t n4 = count / 4;
    for (int64_t i = tid; i < n4; i += stride) {
      const int4 v = d4[i];
      mx = max(mx, max(max(v.x, v.y), max(v.z, v.w)));
    }
    done = n4 * 4;
  }
  for (int64_t i = done + tid; i < count; i += stride) mx = max(mx, d[i]);
  if (__ballot(mx >= cap) != 0ull && (threadIdx.x & 63) == 0) atomicOr(f, 1);
}

}  // namespace

namespace {

__global__ void k_flags_zero2(int* f) { f[threadIdx.x] = 0; }

}  // namespace

int range_check_enqueue(const int32_t* d, int64_t count, int* flags, hipStream_t st) {
  const int64_t work = (count + 3) / 4;
  const int blocks = (int)std::min<int64_t>((work + 255) / 256, 4096);
  const int vec4 = (reinterpret_cast<uintptr_t>(d) & 15u) == 0;
  k_flags_zero2<<<1, 2, 0, st>>>(flags);
  FWX_LAUNCHED("k_flags_zero2");
  k_range<<<blocks > 0 ? blocks : 1, 256, 0, st>>>(d, count, vec4, flags);
  FWX_LAUNCHED("k_range");
  return FWX_OK;
}

int deferred_checks_read(const int32_t* d, int64_t count, int32_t cap, bool scan, int* flags, hipStream_t st,
                         int out[2]) {
  if (scan) {
    const int64_t work = (count + 3) / 4;
    const int blocks = (int)std::min<int64_t>((work + 255) / 256, 4096);
    const int vec4 = (reinterpret_cast<uintptr_t>(d) & 15u) == 0;
    k_at_least<<<blocks > 0 ? blocks : 1, 256, 0, st>>>(d, count, vec4, cap, flags + 1);
    FWX_LAUNCHED("k_at_least");
  }
  hipError_t e = hipMemcpyAsync(out, flags, 2 * sizeof(int), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  return e == hipSuccess ? FWX_OK : hip_fail(e, "flag read-back", __FILE__, __LINE__);
}

int values_below(const int32_t* d, int64_t count, int32_t cap, hipStream_t st, bool* ok) {
  const int64_t work = (count + 3) / 4;
  const int blocks = (int)std::min<int64_t>((work + 255) / 256, 4096);
  const int vec4 = (reinterpret_cast<uintptr_t>(d) & 15u) == 0;
  int h = 1;
  const int rc = flag_query(
      st,
      [&](int* f) -> int {
        k_at_least<<<blocks > 0 ? blocks : 1, 256, 0, st>>>(d, count, vec4, cap, f);
        FWX_LAUNCHED("k_at_least");
        return FWX_OK;
      },
      &h);
  if (rc) return rc;
  *ok = (h == 0);
  return FWX_OK;
}

namespace {

// Any D[i][i] < 0 -> *f = 1.  One diagonal cell per thread.
__global__ __launch_bounds__(256) void k_negative_diagonal(const int32_t* __restrict__ d, int32_t n,
                                                            int* __restrict__ f) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const bool neg = i < n && d[i * n + i] < 0;
  if (__ballot(neg) != 0ull && (threadIdx.x & 63) == 0) atomicOr(f, 1);
}

}  // namespace

int device_negative_cycle(const int32_t* d, int32_t n, hipStream_t st, bool* neg) {
  int h = 0;
  const int rc = flag_query(
      st,
      [&](int* f) -> int {
        k_negative_diagonal<<<(n + 255) / 256, 256, 0, st>>>(d, n, f);
        FWX_LAUNCHED("k_negative_diagonal");
        return FWX_OK;
      },
      &h);
  if (rc) return rc;
  *neg = (h != 0);
  return FWX_OK;
}

}  // namespace fwx
