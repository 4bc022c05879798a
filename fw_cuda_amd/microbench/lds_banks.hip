// lds_banks.hip -- LDS bank-conflict cost of single access patterns on the
// GPU (gfx950): each kernel repeats one LDS access pattern 4096 times per
// wave (256-thread blocks, one grid of 256 blocks), so the rocprofv3
// counters SQ_LDS_BANK_CONFLICT and SQ_LDS_IDX_ACTIVE per kernel, divided by
// the waves and the 4096 repetitions, give the conflict cycles and the LDS
// cycles one instruction of that pattern costs.  The patterns are those of
// the tile product's staging (fwx_tile.hpp): A rows read as 16-byte groups
// (two distinct rows per wave, row stride 36 dwords), B pairs read as 8-byte
// slots (lanes 32-63 repeating lanes 0-31), and the stash writes of both, as
// the 16-byte, 8-byte and paired 4-byte (ds_write2_b32) instructions the
// compiler can emit for them; plus plain 4-, 8- and 16-byte strides to fix
// the bank model.  Every address stays inside the block's 64 KiB array.
// Usage (repository root, `make -C fw_cuda_amd microbench` first):
//   rocprofv3 --pmc SQ_LDS_BANK_CONFLICT SQ_INSTS_LDS SQ_LDS_IDX_ACTIVE SQ_WAVES -- fw_cuda_amd/bin/lds_banks
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdint>
constexpr int ITERS = 4096;

#define KBODY(NAME, SETUP, OP)                                                     \
  __global__ __launch_bounds__(256) void NAME(uint32_t* out) {                     \
    __shared__ __align__(16) uint32_t s[16384];                                    \
    const int l = threadIdx.x;                                                     \
    for (int i = l; i < 16384; i += 256) s[i] = i * 2654435761u;                   \
    __syncthreads();                                                               \
    uint32_t acc = 0;                                                              \
    SETUP                                                                          \
    for (int it = 0; it < ITERS; ++it) {                                           \
      OP                                                                           \
    }                                                                              \
    out[blockIdx.x * 256 + l] = acc + s[l];                                        \
  }

// lane addresses in dwords; (l & 63) is the lane in the wave, w the wave
KBODY(p_b32_s1, const int a = (l & 63) + (l >> 6) * 64;,
      { acc += s[(a + (it & 7) * 256) & 16383]; asm volatile("" : "+v"(acc)); })
KBODY(p_b32_s2, const int a = 2 * (l & 63) + (l >> 6) * 128;,
      { acc += s[(a + (it & 7) * 512) & 16383]; asm volatile("" : "+v"(acc)); })
KBODY(p_b32_s32, const int a = 32 * (l & 63);,
      { acc += s[(a + (it & 7)) & 16383]; asm volatile("" : "+v"(acc)); })
KBODY(p_b32_s64, const int a = 64 * (l & 63) & 16383;,
      { acc += s[(a + (it & 7)) & 16383]; asm volatile("" : "+v"(acc)); })
// b64 consecutive 8-byte slots
KBODY(p_b64_s1, const int a = 2 * (l & 63);,
      { const uint2 v = *reinterpret_cast<const uint2*>(&s[(a + (it & 7) * 128) & 16382]); acc += v.x ^ v.y; asm volatile("" : "+v"(acc)); })
// b64: lanes 32-63 read the same slots as lanes 0-31 (the B operand reads)
KBODY(p_b64_dup, const int a = 2 * (l & 31);,
      { const uint2 v = *reinterpret_cast<const uint2*>(&s[(a + (it & 7) * 128) & 16382]); acc += v.x ^ v.y; asm volatile("" : "+v"(acc)); })
// b64 stride 16 B
KBODY(p_b64_s2, const int a = 4 * (l & 63);,
      { const uint2 v = *reinterpret_cast<const uint2*>(&s[(a + (it & 7) * 256) & 16382]); acc += v.x ^ v.y; asm volatile("" : "+v"(acc)); })
// b128: two distinct rows per wave, row stride 36 dwords (the A operand reads)
KBODY(p_b128_A, const int a = ((l >> 5) & 1) * 36;,
      { const uint4 v = *reinterpret_cast<const uint4*>(&s[(a + (it & 7) * 4) & 16380]); acc += v.x ^ v.y ^ v.z ^ v.w; asm volatile("" : "+v"(acc)); })
// b128 consecutive 16-byte groups (64 distinct)
KBODY(p_b128_s1, const int a = 4 * (l & 63);,
      { const uint4 v = *reinterpret_cast<const uint4*>(&s[(a + (it & 7) * 256) & 16380]); acc += v.x ^ v.y ^ v.z ^ v.w; asm volatile("" : "+v"(acc)); })
// b128 all lanes one address
KBODY(p_b128_bc, const int a = 0;,
      { const uint4 v = *reinterpret_cast<const uint4*>(&s[(a + (it & 7) * 4) & 16380]); acc += v.x ^ v.y ^ v.z ^ v.w; asm volatile("" : "+v"(acc)); })
// write2_b32 pattern of the B stash: lane l writes dwords 2*slot, 2*slot+1,
// slot = l % 32 (+ 128 pair slots for lanes 32-63: the next k-pair row)
KBODY(p_w2_B, const int slot = (l & 31) + ((l >> 5) & 1) * 128 + (l >> 6) * 32;,
      { uint32_t* r = &s[(2 * slot + (it & 7) * 512) & 16382]; r[0] = acc + it; r[1] = acc ^ it; acc += 1; asm volatile("" : "+v"(acc) : : "memory"); })
// ds_write2_b32 (the instruction the kernel's B stash uses) on the same slots
__device__ __forceinline__ void w2(uint32_t* r, uint32_t x, uint32_t y) {
  asm volatile("ds_write2_b32 %0, %1, %2 offset1:1\n s_waitcnt lgkmcnt(0)" : : "v"((uint32_t)(uintptr_t)r), "v"(x), "v"(y) : "memory");
}
KBODY(p_w2asm_B, const int slot = (l & 31) + ((l >> 5) & 1) * 128 + (l >> 6) * 32;,
      { w2(&s[(2 * slot + (it & 7) * 512) & 16382], acc + it, acc ^ it); acc += 1; asm volatile("" : "+v"(acc) : : "memory"); })
// write_b64 of the same slots
KBODY(p_w64_B, const int slot = (l & 31) + ((l >> 5) & 1) * 128 + (l >> 6) * 32;,
      { *reinterpret_cast<uint2*>(&s[(2 * slot + (it & 7) * 512) & 16382]) = make_uint2(acc + it, acc ^ it); acc += 1; asm volatile("" : "+v"(acc) : : "memory"); })
// write_b128 of the A stash: row = tid / 8 (stride 36 dwords), k group = tid % 8
KBODY(p_w128_A, const int a = (l >> 3) * 36 + (l & 7) * 4;,
      { *reinterpret_cast<uint4*>(&s[(a + (it & 3) * 2304) & 16380]) = make_uint4(acc, it, acc, it); acc += 1; asm volatile("" : "+v"(acc) : : "memory"); })
// write_b128 with row stride 32 dwords (no padding)
KBODY(p_w128_A32, const int a = (l >> 3) * 32 + (l & 7) * 4;,
      { *reinterpret_cast<uint4*>(&s[(a + (it & 3) * 2048) & 16380]) = make_uint4(acc, it, acc, it); acc += 1; asm volatile("" : "+v"(acc) : : "memory"); })

int main() {
  uint32_t* out;
  hipMalloc(&out, 256 * 256 * 4);
  void (*ks[])(uint32_t*) = {p_b32_s1, p_b32_s2, p_b32_s32, p_b32_s64, p_b64_s1, p_b64_dup, p_b64_s2, p_b128_A,
                             p_b128_s1, p_b128_bc, p_w2_B, p_w2asm_B, p_w64_B, p_w128_A, p_w128_A32};
  for (auto k : ks) {
    k<<<256, 256>>>(out);
    if (hipDeviceSynchronize() != hipSuccess) { printf("error\n"); return 1; }
  }
  printf("done\n");
  return 0;
}
