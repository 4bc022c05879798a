// pk16_forms.hip -- the packed 16-bit relaxation form on gfx950: semantics of
// v_pk_minimum3_f16 on the bit patterns the form uses, and its throughput.
//
// pk16 form.  A cell value v in [0, CAP] is held in a 16-bit lane.  Four sums
// come from ONE 64-bit add (v_lshl_add_u64) of
//   a = (A[i][k0]+BIAS, A[i][k0]+BIAS, A[i][k1]+BIAS, A[i][k1]+BIAS)
//   b = (B[k0][j0],     B[k0][j1],     B[k1][j0],     B[k1][j1])
// (16-bit lanes, low to high; no lane carries since every lane sum is below
// 2^16), and ONE v_pk_minimum3_f16 folds them into the two cells (j0, j1) of
// row i:  c = min3(c, sums(k0), sums(k1)) per lane.  Every compared pattern
// (c and the sums) lies in [BIAS, 0x7BFF] = [0x0400, 0x7BFF]: positive normal
// binary16 numbers, which order exactly as their unsigned bit patterns do,
// so the f16 minimum is the unsigned minimum.  2 VALU ops per 4 relaxations.
//
// Part 1 checks that claim on the device: every pattern x in [0, 0x7BFF] in
// either lane against hashed partners (plus x itself and x +- 1), counting
// results that differ from the unsigned per-lane minimum, separately for the
// normal range the form uses and for the subnormal range it avoids; then the
// whole relaxation (64-bit add + min3) against plain integer arithmetic on
// hashed operands in the form's ranges.  Part 2 times the form like
// valu_forms.hip times the others (8 x 8 register block, operands through an
// empty asm each iteration).  Usage: pk16_forms (no arguments).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

#include "../csrc/fwx_device.hpp"

using namespace fwx;

#define OPAQUE(x) asm volatile("" : "+v"(x))

namespace {

constexpr uint32_t kBias = 0x0400u;
constexpr uint32_t kCap = 0x3BFFu;

__device__ __forceinline__ uint32_t hash32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

__device__ __forceinline__ uint32_t pkmin3(uint32_t a, uint32_t b, uint32_t c) {
  uint32_t r;
  asm volatile("v_pk_minimum3_f16 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
  return r;
}

__device__ __forceinline__ uint32_t umin16x2(uint32_t a, uint32_t b) {
  const uint32_t lo = min(a & 0xFFFFu, b & 0xFFFFu);
  const uint32_t hi = min(a >> 16, b >> 16);
  return lo | (hi << 16);
}

// thread t: pattern x = t % 0x7C00 in lane (t / 0x7C00) % 2, variant t / 0xF800
// (0..7): partners hashed, or x itself / x + 1 / x - 1 (clamped to the range).
// cnt[0]: mismatches where all three lane inputs are >= BIAS, cnt[1]: others.
__global__ __launch_bounds__(256) void k_min3_semantics(uint32_t* cnt, uint32_t total) {
  const uint32_t t = blockIdx.x * 256u + threadIdx.x;
  if (t >= total) return;
  const uint32_t x = t % 0x7C00u;
  const uint32_t lane = (t / 0x7C00u) & 1u;
  const uint32_t var = t / 0xF800u;
  const uint32_t h1 = hash32(t * 3u + 1u), h2 = hash32(t * 3u + 2u), h3 = hash32(t * 3u + 3u);
  uint32_t y = h1 % 0x7C00u, z = h2 % 0x7C00u, w = h3 % 0x7C00u;  // other-lane partners too
  if (var == 1) y = x;
  if (var == 2) y = x + 1u < 0x7C00u ? x + 1u : x;
  if (var == 3) z = x > 0u ? x - 1u : x;
  if (var == 4) { y = x; z = x; }
  // lane `lane` carries (x, y, z); the other lane hashed partners; every
  // pattern is below 0x7C00 (finite)
  const uint32_t oy = (h1 >> 16) % 0x7C00u, oz = (h2 >> 16) % 0x7C00u;
  const uint32_t xs2 = lane ? ((x << 16) | w) : (x | (w << 16));
  const uint32_t ys2 = lane ? ((y << 16) | oy) : (y | (oy << 16));
  const uint32_t zs2 = lane ? ((z << 16) | oz) : (z | (oz << 16));
  const uint32_t got = pkmin3(xs2, ys2, zs2);
  const uint32_t want = umin16x2(xs2, umin16x2(ys2, zs2));
  for (int l = 0; l < 2; ++l) {
    const uint32_t g = (got >> (16 * l)) & 0xFFFFu, e = (want >> (16 * l)) & 0xFFFFu;
    const bool normal = ((xs2 >> (16 * l)) & 0xFFFFu) >= kBias && ((ys2 >> (16 * l)) & 0xFFFFu) >= kBias &&
                        ((zs2 >> (16 * l)) & 0xFFFFu) >= kBias;
    if (g != e) atomicAdd(&cnt[normal ? 0 : 1], 1u);
  }
}

// the whole pk16 relaxation on hashed operands in the form's ranges:
// A values in [0, CAP] (+BIAS), B values in [0, CAP], C in [0, CAP] (+BIAS).
__global__ __launch_bounds__(256) void k_relax_semantics(uint32_t* cnt, uint32_t total) {
  const uint32_t t = blockIdx.x * 256u + threadIdx.x;
  if (t >= total) return;
  uint32_t v[8];
  for (int i = 0; i < 8; ++i) {
    const uint32_t h = hash32(t * 8u + i);
    // mostly small values (the common case), sometimes anywhere up to CAP
    v[i] = (h & 0x80000000u) ? (h % (kCap + 1u)) : (h % 64u);
  }
  const uint32_t a0 = v[0] + kBias, a1 = v[1] + kBias;
  const uint64_t a = (uint64_t)(a0 | (a0 << 16)) | ((uint64_t)(a1 | (a1 << 16)) << 32);
  const uint64_t b = (uint64_t)(v[2] | (v[3] << 16)) | ((uint64_t)(v[4] | (v[5] << 16)) << 32);
  const uint32_t c = (v[6] + kBias) | ((v[7] + kBias) << 16);
  const uint64_t s = a + b;
  const uint32_t got = pkmin3(c, (uint32_t)s, (uint32_t)(s >> 32));
  const uint32_t e0 = min(v[6], min(v[0] + v[2], v[1] + v[4])) + kBias;
  const uint32_t e1 = min(v[7], min(v[0] + v[3], v[1] + v[5])) + kBias;
  if (got != (e0 | (e1 << 16))) atomicAdd(&cnt[2], 1u);
}

__device__ __forceinline__ uint32_t relax4(uint32_t c, uint64_t a, uint64_t b) {
  const uint64_t s = a + b;
  return pkmin3(c, (uint32_t)s, (uint32_t)(s >> 32));
}

// 8 x 8 block: c[i][j] holds two cells, each pair (a[i], b[j]) is 4 relaxations
__global__ __launch_bounds__(256) void k_pk16(const uint32_t* in, uint32_t* out, int iters) {
  uint64_t a[8], b[8];
  uint32_t c[64];
  const int t = threadIdx.x;
  for (int i = 0; i < 8; ++i) {
    const uint32_t x = (in[(t + i) & 63] & 0xFFu) + kBias;
    const uint32_t y = (in[(t + i + 5) & 63] & 0xFFu) + kBias;
    a[i] = (uint64_t)(x | (x << 16)) | ((uint64_t)(y | (y << 16)) << 32);
    b[i] = (uint64_t)(in[(t + 3 * i + 1) & 63] & 0x00FF00FFu) | ((uint64_t)(in[(t + 3 * i + 2) & 63] & 0x00FF00FFu) << 32);
  }
  for (int i = 0; i < 64; ++i) c[i] = 0x3F003F00u;
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      OPAQUE(a[i]);
      OPAQUE(b[i]);
    }
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
      for (int j = 0; j < 8; ++j) c[i * 8 + j] = relax4(c[i * 8 + j], a[i], b[j]);
  }
  uint32_t s = 0;
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
  uint32_t* cnt = nullptr;
  uint32_t* in = nullptr;
  uint32_t* out = nullptr;
  const int max_blocks = p.multiProcessorCount * 16;
  uint32_t h[64];
  for (int i = 0; i < 64; ++i) h[i] = (uint32_t)(i * 2654435761u);
  if (!ok(hipMalloc(&cnt, 4 * sizeof(uint32_t)), "hipMalloc") || !ok(hipMemset(cnt, 0, 4 * sizeof(uint32_t)), "hipMemset") ||
      !ok(hipMalloc(&in, sizeof(h)), "hipMalloc") ||
      !ok(hipMalloc(&out, (size_t)max_blocks * 256 * sizeof(uint32_t)), "hipMalloc") ||
      !ok(hipMemcpy(in, h, sizeof(h), hipMemcpyHostToDevice), "hipMemcpy"))
    return 1;
  // part 1: semantics
  const uint32_t total1 = 0xF800u * 8u;  // 8 variants x both lanes x every pattern < 0x7C00
  k_min3_semantics<<<(total1 + 255) / 256, 256>>>(cnt, total1);
  const uint32_t total2 = 1u << 24;
  k_relax_semantics<<<total2 / 256, 256>>>(cnt, total2);
  uint32_t hc[4];
  if (!ok(hipGetLastError(), "launch") || !ok(hipMemcpy(hc, cnt, sizeof(hc), hipMemcpyDeviceToHost), "hipMemcpy"))
    return 1;
  std::printf("v_pk_minimum3_f16 vs unsigned min: %u lane trials, mismatches: normal range %u, with subnormal inputs %u\n",
              total1 * 2u, hc[0], hc[1]);
  std::printf("pk16 relaxation vs integer arithmetic: %u trials, mismatches %u\n", total2, hc[2]);
  // part 2: throughput
  hipEvent_t e0, e1;
  if (!ok(hipEventCreate(&e0), "hipEventCreate") || !ok(hipEventCreate(&e1), "hipEventCreate")) return 1;
  const int iters = 4096;
  const double peak2 = p.multiProcessorCount * 64.0 * p.clockRate * 1e3 / 2.0;
  for (int occ = 0; occ < 2; ++occ) {
    const int nb = occ == 0 ? max_blocks : p.multiProcessorCount * 2;
    float ms = 0.f;
    for (int rep = 0; rep < 2; ++rep) {
      if (!ok(hipEventRecord(e0), "hipEventRecord")) return 1;
      k_pk16<<<nb, 256>>>(in, out, iters);
      if (!ok(hipGetLastError(), "launch") || !ok(hipEventRecord(e1), "hipEventRecord") ||
          !ok(hipEventSynchronize(e1), "hipEventSynchronize") || !ok(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime"))
        return 1;
    }
    const double relax = 256.0 * iters * nb * 256.0;
    const double rate = relax / (ms * 1e-3);
    std::printf("form=pk16     blocks=%5d ms=%8.3f relax/s=%.4e ratio_to_2op_peak=%.3f\n", nb, ms, rate, rate / peak2);
  }
  const hipError_t err = hipDeviceSynchronize();
  std::printf("status=%s\n", hipGetErrorString(err));
  return (err == hipSuccess && hc[0] == 0 && hc[2] == 0) ? 0 : 1;
}
