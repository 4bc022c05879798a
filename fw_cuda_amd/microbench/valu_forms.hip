// valu_forms.hip -- VALU throughput of the three min-plus relaxation forms on
// the GPU (gfx950), measured in isolation from memory: every thread keeps an
// 8 x 8 block of accumulators and 8 + 8 operands in registers and relaxes it
// `iters` times.
//
//   add+min : c = min(c, a + b)                       2 ops / relaxation
//   min3    : c = v_min3_i32(c, a0 + b0, a1 + b1)     3 ops / 2 relaxations
//   pk64    : s = (a1:a0) + (b1:b0) as one 64-bit add, c = v_min3_i32(c, s.lo, s.hi)
//                                                     2 ops / 2 relaxations
//
// The min3 and pk64 forms are the library's own (fwx_device.hpp).  The
// operands pass through an empty asm each iteration so the compiler cannot
// hoist the relaxations out of the loop.  Printed: relaxations/s and the
// ratio to the 2-op peak CUs x 64 lanes x clock / 2 at the clock the device
// reports.  Usage: valu_forms (no arguments).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

#include "../csrc/fwx_device.hpp"

using namespace fwx;

#define OPAQUE(x) asm volatile("" : "+v"(x))

namespace {

constexpr int kOps = 64;  // input words, indexed & 63

__global__ __launch_bounds__(256) void k_addmin(const int32_t* in, int32_t* out, int iters) {
  int32_t a[8], b[8], c[64];
  const int t = threadIdx.x;
  for (int i = 0; i < 8; ++i) {
    a[i] = in[(t + i) & (kOps - 1)];
    b[i] = in[(t + 3 * i + 1) & (kOps - 1)];
  }
  for (int i = 0; i < 64; ++i) c[i] = in[(t + i + 7) & (kOps - 1)] + 1000000;
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      OPAQUE(a[i]);
      OPAQUE(b[i]);
    }
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
      for (int j = 0; j < 8; ++j) c[i * 8 + j] = min(c[i * 8 + j], a[i] + b[j]);
  }
  int32_t s = 0;
#pragma unroll
  for (int i = 0; i < 64; ++i) s ^= c[i];
  out[blockIdx.x * blockDim.x + t] = s;
}

template <bool PK>
__global__ __launch_bounds__(256) void k_pair(const int32_t* in, int32_t* out, int iters) {
  uint64_t a[8], b[8];
  int32_t c[64];
  const int t = threadIdx.x;
  for (int i = 0; i < 8; ++i) {
    a[i] = pack2(in[(t + i) & (kOps - 1)], in[(t + i + 5) & (kOps - 1)]);
    b[i] = pack2(in[(t + 3 * i + 1) & (kOps - 1)], in[(t + 3 * i + 2) & (kOps - 1)]);
  }
  for (int i = 0; i < 64; ++i) c[i] = in[(t + i + 7) & (kOps - 1)] + 1000000;
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      OPAQUE(a[i]);
      OPAQUE(b[i]);
    }
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
      for (int j = 0; j < 8; ++j)
        c[i * 8 + j] = PK ? relax2_pk64(c[i * 8 + j], a[i], b[j]) : relax2_min3(c[i * 8 + j], a[i], b[j]);
  }
  int32_t s = 0;
#pragma unroll
  for (int i = 0; i < 64; ++i) s ^= c[i];
  out[blockIdx.x * blockDim.x + t] = s;
}

bool ok(hipError_t e, const char* what) {
  if (e == hipSuccess) return true;
  std::fprintf(stderr, "%s: %s\n", what, hipGetErrorString(e));
  return false;
}

}  // namespace

int main() {
  hipDeviceProp_t p;
  if (!ok(hipGetDeviceProperties(&p, 0), "hipGetDeviceProperties")) return 1;
  std::printf("device %s (%s), %d CUs, clock %.0f MHz\n", p.name, p.gcnArchName, p.multiProcessorCount,
              p.clockRate / 1e3);
  int32_t* in = nullptr;
  int32_t* out = nullptr;
  const int max_blocks = p.multiProcessorCount * 16;  // 16 blocks of 4 waves per CU
  int32_t h[kOps];
  for (int i = 0; i < kOps; ++i) h[i] = (i * 7919) % 1000 + 1;
  if (!ok(hipMalloc(&in, sizeof(h)), "hipMalloc") ||
      !ok(hipMalloc(&out, (size_t)max_blocks * 256 * sizeof(int32_t)), "hipMalloc") ||
      !ok(hipMemcpy(in, h, sizeof(h), hipMemcpyHostToDevice), "hipMemcpy"))
    return 1;
  hipEvent_t e0, e1;
  if (!ok(hipEventCreate(&e0), "hipEventCreate") || !ok(hipEventCreate(&e1), "hipEventCreate")) return 1;
  const int iters = 4096;
  const double peak2 = p.multiProcessorCount * 64.0 * p.clockRate * 1e3 / 2.0;
  const char* names[3] = {"add+min", "min3", "pk64"};
  for (int occ = 0; occ < 2; ++occ) {
    const int nb = occ == 0 ? max_blocks : p.multiProcessorCount * 2;
    for (int kind = 0; kind < 3; ++kind) {
      float ms = 0.f;
      for (int rep = 0; rep < 2; ++rep) {  // first run warms up
        if (!ok(hipEventRecord(e0), "hipEventRecord")) return 1;
        if (kind == 0) k_addmin<<<nb, 256>>>(in, out, iters);
        if (kind == 1) k_pair<false><<<nb, 256>>>(in, out, iters);
        if (kind == 2) k_pair<true><<<nb, 256>>>(in, out, iters);
        if (!ok(hipGetLastError(), "launch") || !ok(hipEventRecord(e1), "hipEventRecord") ||
            !ok(hipEventSynchronize(e1), "hipEventSynchronize") ||
            !ok(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime"))
          return 1;
      }
      const double relax = (kind == 0 ? 64.0 : 128.0) * iters * nb * 256.0;
      const double rate = relax / (ms * 1e-3);
      std::printf("form=%-8s blocks=%5d waves/SIMD=%2d ms=%8.3f relax/s=%.4e ratio_to_2op_peak=%.3f\n", names[kind],
                  nb, nb * 4 / (p.multiProcessorCount * 4), ms, rate, rate / peak2);
    }
  }
  const hipError_t err = hipDeviceSynchronize();
  std::printf("status=%s\n", hipGetErrorString(err));
  return err == hipSuccess ? 0 : 1;
}
