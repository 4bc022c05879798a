// fwx_pred.hip -- predecessor (path-reconstruction) matrix for the int32
// engine.  No reference counterpart: the reference returns distances only.
//
// Given the input weights W (INF-completed, as the reference treats them)
// and the solved distances D = closure(W):
//   P[i][j] = the smallest u != j with D[i][u] + W[u][j] == D[i][j],
//   P[i][i] = -1, and -1 where no u matches (D not the closure of W).
// Because D is closed, D[i][u] + W[u][j] >= D[i][j] for every u, so the
// matches are exactly the minimisers: each cell is a min over
//   key(u) = (min(D[i][u] + W[u][j] - D[i][j], 0x7FFF) << 16) | u,
// and a result with a zero upper half is a match.  When every cycle of the
// graph has positive weight, walking P from j reaches i along a shortest
// path (each step strictly lowers D[i][.] or follows a zero-weight edge into
// a vertex closer in the tight subgraph, which is acyclic then).
//
// One n^3 pass, tiled like the products of the solver: 64 x 64 P blocks of
// 256 threads (4 x 4 cells per thread), u walked in chunks of 32 through
// LDS.  u == j is excluded by staging W[j][j] as INF.
#include <hip/hip_runtime.h>

#include "fwx_device.hpp"
#include "fwx_internal.hpp"

namespace fwx {

namespace {

constexpr int kPT = 64;   // P tile
constexpr int kPU = 32;   // u per chunk
constexpr int kPLA = kPU + 4;  // A row stride (dwords)

__global__ __launch_bounds__(256) void k_predecessors(const int32_t* __restrict__ W, const int32_t* __restrict__ D,
                                                       int32_t* __restrict__ P, int32_t n) {
  __shared__ int32_t As[kPT * kPLA];   // D[i][u], row-major
  __shared__ int32_t Bs[kPU * kPT];    // W[u][j]
  const int tid = threadIdx.x;
  const int tx = tid & 15, ty = tid >> 4;
  const int i0 = blockIdx.y * kPT, j0 = blockIdx.x * kPT;
  int32_t nd[4][4];
  uint32_t best[4][4];
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int i = i0 + ty + 16 * r, j = j0 + tx + 16 * c;
      nd[r][c] = (i < n && j < n) ? -D[(int64_t)i * n + j] : 0;
      best[r][c] = 0x7FFFFFFFu;
    }
  for (int u0 = 0; u0 < n; u0 += kPU) {
    // stage A: 64 rows x 32 u of D (2048 cells, 8 per thread)
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int idx = e * 256 + tid;
      const int r = idx / kPU, u = idx % kPU;
      const int i = i0 + r, uu = u0 + u;
      As[r * kPLA + u] = (i < n && uu < n) ? D[(int64_t)i * n + uu] : kInf;
    }
    // stage B: 32 u x 64 columns of W, W[j][j] -> INF
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int idx = e * 256 + tid;
      const int u = idx / kPT, c = idx % kPT;
      const int uu = u0 + u, j = j0 + c;
      Bs[u * kPT + c] = (uu < n && j < n && uu != j) ? W[(int64_t)uu * n + j] : kInf;
    }
    __syncthreads();
#pragma unroll 2
    for (int u = 0; u < kPU; u += 2) {
      int32_t a0[4], a1[4], b0[4], b1[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        a0[r] = As[(ty + 16 * r) * kPLA + u];
        a1[r] = As[(ty + 16 * r) * kPLA + u + 1];
      }
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        b0[c] = Bs[u * kPT + tx + 16 * c];
        b1[c] = Bs[(u + 1) * kPT + tx + 16 * c];
      }
      const uint32_t uk0 = (uint32_t)(u0 + u), uk1 = uk0 + 1u;
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const uint32_t t0 = (uint32_t)a0[r] + (uint32_t)b0[c] + (uint32_t)nd[r][c];
          const uint32_t t1 = (uint32_t)a1[r] + (uint32_t)b1[c] + (uint32_t)nd[r][c];
          const uint32_t k0 = (min(t0, 0x7FFFu) << 16) | uk0;
          const uint32_t k1 = (min(t1, 0x7FFFu) << 16) | uk1;
          best[r][c] = min(best[r][c], min(k0, k1));
        }
    }
    __syncthreads();
  }
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int i = i0 + ty + 16 * r, j = j0 + tx + 16 * c;
      if (i < n && j < n)
        P[(int64_t)i * n + j] = (i == j || (best[r][c] >> 16) != 0u) ? -1 : (int32_t)(best[r][c] & 0xFFFFu);
    }
}

}  // namespace

}  // namespace fwx

using namespace fwx;

extern "C" {

int fwx_predecessors(const int32_t* d_W, const int32_t* d_D, int32_t* d_P, int32_t n, void* stream) {
  clear_error();
  if (!d_W || !d_D || !d_P || n <= 0 || n > FWX_MAX_N)
    return fail(FWX_ERR_INVALID_ARG, "fwx_predecessors: null pointer or n out of range");
  const int nt = (n + kPT - 1) / kPT;
  hipStream_t st = as_stream(stream);
  k_predecessors<<<dim3(nt, nt), 256, 0, st>>>(d_W, d_D, d_P, n);
  FWX_LAUNCHED("k_predecessors");
  FWX_HIP(hipStreamSynchronize(st));
  return FWX_OK;
}

}  // extern "C"
