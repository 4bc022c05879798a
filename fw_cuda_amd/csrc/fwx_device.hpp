// fwx_device.hpp -- gfx950 building blocks shared by the Floyd-Warshall kernels.
//
// A relaxation is c = min(c, a + b) on int32 with two's-complement wrap on
// the add and a signed compare, the arithmetic of the reference's
// `D[i*n+k] + D[k*n+j] < D[i*n+j]` (src/fw_baseline_cpu.cpp:9-12).
//
// Two VALU forms of the same arithmetic are used on a pair of k values
// (k0, k1) feeding one output cell:
//   * min3 form:  c = v_min3_i32(c, a0 + b0, a1 + b1)        3 ops / 2 relax
//   * pk64 form:  s = v_lshl_add_u64((a1:a0), 0, (b1:b0))     1 op
//                 c = v_min3_i32(c, s.lo, s.hi)                1 op
//     The 64-bit add produces both 32-bit sums in one instruction.  It is
//     exact when the low half cannot carry into the high half, i.e. when
//     a0 + b0 < 2^32 as unsigned; every operand in [0, 2^30) guarantees it
//     (and then no int32 sum overflows either, so the signed compare sees the
//     same value as the reference).  If every input cell lies in [0, 2^30),
//     so does every value a Floyd-Warshall run produces (each is a sum of
//     input cells along a walk, or an input cell, and only ever decreases),
//     so one check of the input per call (fwx_range.hip) selects the form for
//     the whole run.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fwx {

constexpr int32_t kInf = 1000000000;  // include/constants.h:3
// A value with either of the top two bits set rules out the pk64 form.
constexpr uint32_t kPk64Mask = 0xC0000000u;

__device__ __forceinline__ int32_t wadd(int32_t a, int32_t b) {
  return (int32_t)((uint32_t)a + (uint32_t)b);
}

__device__ __forceinline__ int32_t min3i(int32_t a, int32_t b, int32_t c) {
  return min(a, min(b, c));
}

// v_min3_i32 as one instruction.  Written as asm because the compiler
// otherwise re-associates chains of min3 into min trees (3 min ops per two
// pairs instead of 2).  Pure ALU, no memory operands.
__device__ __forceinline__ int32_t vmin3(int32_t a, int32_t b, int32_t c) {
  int32_t r;
  asm("v_min3_i32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
  return r;
}

// v_min3_u32 as one instruction (same reason as vmin3).
__device__ __forceinline__ uint32_t vmin3u(uint32_t a, uint32_t b, uint32_t c) {
  uint32_t r;
  asm("v_min3_u32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
  return r;
}

// pair relaxation in pk64 form: a = (a1:a0), b = (b1:b0)
__device__ __forceinline__ int32_t relax2_pk64(int32_t c, uint64_t a, uint64_t b) {
  const uint64_t s = a + b;
  return vmin3(c, (int32_t)(uint32_t)s, (int32_t)(uint32_t)(s >> 32));
}

// pair relaxation in min3 form
__device__ __forceinline__ int32_t relax2_min3(int32_t c, uint64_t a, uint64_t b) {
  const int32_t s0 = wadd((int32_t)(uint32_t)a, (int32_t)(uint32_t)b);
  const int32_t s1 = wadd((int32_t)(uint32_t)(a >> 32), (int32_t)(uint32_t)(b >> 32));
  return vmin3(c, s0, s1);
}

__device__ __forceinline__ uint64_t pack2(int32_t lo, int32_t hi) {
  return (uint64_t)(uint32_t)lo | ((uint64_t)(uint32_t)hi << 32);
}

}  // namespace fwx
