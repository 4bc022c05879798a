// fwx_gen.hip -- graphs generated on the GPU, and the pinned / asynchronous
// transfers around the solve.
//
// The reference generates its inputs on the host, one std::mt19937 draw
// after another (src/graph_generator.cpp:25-43: ~2e9 draws at n=32768), then
// copies them with a pageable cudaMemcpy (src/gpu_memory.cu:13-16).  The
// device generator below draws every cell independently from a
// counter-based generator, Philox4x32-10 (Salmon, Moraes, Dror, Shaw,
// "Parallel random numbers: as easy as 1, 2, 3", SC'11): the 128-bit counter
// is the cell index, the 64-bit key the seed.  Same distributions as the
// reference -- an edge with probability `density`, weight uniform in
// [wmin, wmax], diagonal 0, INF where there is no edge -- but NOT the same
// stream: a seed gives a different matrix than fwx_generate_graph.  Every
// cell depends only on (seed, i, j), so the output is identical for any
// launch shape and on any device.
//
// Per cell: (c0, c1, c2, c3) = Philox(counter = (cell lo, cell hi, 0, 0),
// key = (seed lo, seed hi)); present iff c0 < floor(density * 2^32) (always
// for density 1); weight = wmin + floor(c1 * (wmax - wmin + 1) / 2^32).
#include <hip/hip_runtime.h>

#include "fwx_device.hpp"
#include "fwx_internal.hpp"

namespace fwx {

namespace {

struct Philox4 {
  uint32_t x[4];
};

// Philox4x32 with 10 rounds: multipliers 0xD2511F53 / 0xCD9E8D57, key
// increments 0x9E3779B9 / 0xBB67AE85 (the published constants).
__device__ __forceinline__ Philox4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                                                  uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    c1 = (uint32_t)p1;
    c3 = (uint32_t)p0;
    c0 = n0;
    c2 = n2;
  }
  return Philox4{{c0, c1, c2, c3}};
}

__device__ __forceinline__ int32_t gen_cell(uint64_t cell, bool diag, uint64_t thresh, uint32_t k0, uint32_t k1,
                                            int32_t wmin, uint64_t range) {
  if (diag) return 0;
  const Philox4 r = philox4x32_10((uint32_t)cell, (uint32_t)(cell >> 32), 0u, 0u, k0, k1);
  if ((uint64_t)r.x[0] >= thresh) return kInf;
  return wmin + (int32_t)(((uint64_t)r.x[1] * range) >> 32);
}

// One thread per 4 consecutive cells of a row (16-byte store when n % 4 == 0
// and the buffer is 16-byte aligned).
template <bool VEC>
__global__ __launch_bounds__(256) void k_generate(int32_t* __restrict__ d, int32_t n, uint64_t thresh, uint32_t k0,
                                                   uint32_t k1, int32_t wmin, uint64_t range) {
  const int32_t q = (n + 3) / 4;  // 4-cell groups per row
  const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (g >= (int64_t)n * q) return;
  const int32_t i = (int32_t)(g / q);
  const int32_t j0 = (int32_t)(g % q) * 4;
  const uint64_t base = (uint64_t)i * (uint64_t)n;
  int32_t v[4];
#pragma unroll
  for (int e = 0; e < 4; ++e)
    v[e] = j0 + e < n ? gen_cell(base + (uint64_t)(j0 + e), i == j0 + e, thresh, k0, k1, wmin, range) : 0;
  if (VEC) {
    *reinterpret_cast<int4*>(d + base + j0) = make_int4(v[0], v[1], v[2], v[3]);
  } else {
#pragma unroll
    for (int e = 0; e < 4; ++e)
      if (j0 + e < n) d[base + (uint64_t)(j0 + e)] = v[e];
  }
}

}  // namespace

}  // namespace fwx

using namespace fwx;

extern "C" {

int fwx_generate_graph_device(int32_t* d_D, int32_t n, double density, uint64_t seed, int32_t wmin, int32_t wmax,
                              void* stream) {
  clear_error();
  if (!d_D || n <= 0 || n > FWX_MAX_N) return fail(FWX_ERR_INVALID_ARG, "null pointer or n out of range");
  if (!(density >= 0.0 && density <= 1.0)) return fail(FWX_ERR_INVALID_ARG, "density must be in [0, 1]");
  if (wmin > wmax) return fail(FWX_ERR_INVALID_ARG, "wmin must be <= wmax");
  const uint64_t thresh = density >= 1.0 ? (1ull << 32) : (uint64_t)(density * 4294967296.0);
  const uint64_t range = (uint64_t)((int64_t)wmax - (int64_t)wmin + 1);
  const int64_t groups = (int64_t)n * ((n + 3) / 4);
  const int blocks = (int)((groups + 255) / 256);
  hipStream_t st = as_stream(stream);
  const bool vec = (n % 4 == 0) && ((reinterpret_cast<uintptr_t>(d_D) & 15u) == 0);
  if (vec)
    k_generate<true><<<blocks, 256, 0, st>>>(d_D, n, thresh, (uint32_t)seed, (uint32_t)(seed >> 32), wmin, range);
  else
    k_generate<false><<<blocks, 256, 0, st>>>(d_D, n, thresh, (uint32_t)seed, (uint32_t)(seed >> 32), wmin, range);
  FWX_LAUNCHED("k_generate");
  FWX_HIP(hipStreamSynchronize(st));
  return FWX_OK;
}

int fwx_host_alloc_pinned(int32_t** h_D, int32_t n) {
  clear_error();
  if (!h_D) return fail(FWX_ERR_INVALID_ARG, "null output pointer");
  if (n <= 0 || n > FWX_MAX_N) return fail(FWX_ERR_INVALID_ARG, "n must be in [1, 32769]");
  *h_D = nullptr;
  FWX_HIP(hipHostMalloc(reinterpret_cast<void**>(h_D), (size_t)n * (size_t)n * sizeof(int32_t), hipHostMallocDefault));
  return FWX_OK;
}

int fwx_host_free_pinned(int32_t* h_D) {
  clear_error();
  if (h_D) FWX_HIP(hipHostFree(h_D));
  return FWX_OK;
}

int fwx_copy_to_device_async(int32_t* d_D, const int32_t* h_D, int32_t n, void* stream) {
  clear_error();
  if (!d_D || !h_D || n <= 0 || n > FWX_MAX_N) return fail(FWX_ERR_INVALID_ARG, "null pointer or n out of range");
  FWX_HIP(hipMemcpyAsync(d_D, h_D, (size_t)n * (size_t)n * sizeof(int32_t), hipMemcpyHostToDevice, as_stream(stream)));
  return FWX_OK;
}

int fwx_copy_to_host_async(int32_t* h_D, const int32_t* d_D, int32_t n, void* stream) {
  clear_error();
  if (!d_D || !h_D || n <= 0 || n > FWX_MAX_N) return fail(FWX_ERR_INVALID_ARG, "null pointer or n out of range");
  FWX_HIP(hipMemcpyAsync(h_D, d_D, (size_t)n * (size_t)n * sizeof(int32_t), hipMemcpyDeviceToHost, as_stream(stream)));
  return FWX_OK;
}

int fwx_stream_synchronize(void* stream) {
  clear_error();
  FWX_HIP(hipStreamSynchronize(as_stream(stream)));
  return FWX_OK;
}

}  // extern "C"
