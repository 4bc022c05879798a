// fwx_baseline.hip -- baseline_gpu: one HBM-bound relaxation sweep per k.
//
// Replaces fwBaselineGPU (src/fw_baseline_gpu.cu:16-32), which launches one
// kernel per k computing D[i][j] = min(D[i][j], D[i][k] + D[k][j]) over the
// whole matrix.  Same per-k update and launch structure; each thread owns a
// 4-column group of one row (16-byte loads when n % 4 == 0) and stores only
// when a cell improved.  Within step k, row k and column k are unchanged
// whenever D[k][k] >= 0, so threads need no ordering among themselves (the
// reference relies on the same property).
#include <hip/hip_runtime.h>

#include "fwx_device.hpp"
#include "fwx_internal.hpp"

namespace fwx {

template <bool VEC>
__global__ __launch_bounds__(256) void k_baseline_step(int32_t* __restrict__ D, int32_t n,
                                                       int32_t k) {
  const int32_t i = blockIdx.y * 4 + threadIdx.y;
  const int32_t j = (blockIdx.x * 64 + threadIdx.x) * 4;
  if (i >= n || j >= n) return;
  int32_t* row = D + (int64_t)i * n;
  const int32_t* rowk = D + (int64_t)k * n;
  const int32_t dik = row[k];
  if (VEC) {
    int4 c = *reinterpret_cast<const int4*>(row + j);
    const int4 b = *reinterpret_cast<const int4*>(rowk + j);
    const int4 o = c;
    c.x = min(c.x, wadd(dik, b.x));
    c.y = min(c.y, wadd(dik, b.y));
    c.z = min(c.z, wadd(dik, b.z));
    c.w = min(c.w, wadd(dik, b.w));
    if (c.x != o.x || c.y != o.y || c.z != o.z || c.w != o.w) *reinterpret_cast<int4*>(row + j) = c;
  } else {
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      if (j + e < n) {
        const int32_t v = wadd(dik, rowk[j + e]);
        if (v < row[j + e]) row[j + e] = v;
      }
    }
  }
}

int enqueue_baseline(int32_t* d, int32_t n, hipStream_t st) {
  const bool vec = (n % 4 == 0) && ((reinterpret_cast<uintptr_t>(d) & 15u) == 0);
  const dim3 block(64, 4);
  const dim3 grid((n + 255) / 256, (n + 3) / 4);
  const double bytes = 4.0 * (double)n * (double)n;  // every cell is read once per k
  // While profiling, one event per sweep: each ends the previous sweep's
  // group and starts the next (the group time is launch to launch).
  hipEvent_t ev;
  if (int rc = prof_open(st, &ev)) return rc;
  for (int32_t k = 0; k < n; ++k) {
    if (vec)
      k_baseline_step<true><<<grid, block, 0, st>>>(d, n, k);
    else
      k_baseline_step<false><<<grid, block, 0, st>>>(d, n, k);
    FWX_LAUNCHED("k_baseline_step");
    if (int rc = prof_chain(st, &ev, FWX_PROF_BASELINE, bytes)) return rc;
  }
  return FWX_OK;
}

}  // namespace fwx
