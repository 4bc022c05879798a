// fwx_pred.hip -- predecessor (path-reconstruction) matrix for the int32
// engine.  No reference counterpart: the reference returns distances only.
//
// Given the input weights W (INF-completed, as the reference treats them)
// and the solved distances D = closure(W):
//   P[i][j] = the smallest u != j with D[i][u] + W[u][j] == D[i][j],
//   P[i][i] = -1, and -1 where no u matches (D not the closure of W).
// Because D is closed, D[i][u] + W[u][j] >= D[i][j] for every u, so the
// matches are exactly the minimisers: each cell is a min over
//   key(u) = (min(D[i][u] + W[u][j] - D[i][j], 0xFFFF) << 16) | u,
// and a result with a zero upper half is a match.  When every cycle of the
// graph has positive weight, walking P from j reaches i along a shortest
// path (each step strictly lowers D[i][.] or follows a zero-weight edge into
// a vertex closer in the tight subgraph, which is acyclic then).
//
// One n^3 pass, tiled like the products of the solver: 64 x 64 P blocks of
// 256 threads, each thread 4 rows (ty + 16 r) x 4 adjacent columns
// (4 tx + c), u walked in chunks of 32 through LDS and 4 at a time in
// registers (16-byte LDS reads of A rows and B rows).  u == j is excluded by
// staging W[j][j] as INF.  Per (i, j, u): t = D[i][u] + W[u][j] - D[i][j]
// (v_add3_u32; t >= 0 as D is closed, and any negative t of a malformed D
// wraps above 2^31 and saturates like a large one), the key
// (sat16(t) << 16) | u in one v_cvt_pk_u16_u32 (it clamps each 32-bit
// operand to 16 bits; u < 2^16), and one v_min3_u32 per two keys: 2.5 VALU
// operations per relaxation.
#include <hip/hip_runtime.h>

#include "fwx_device.hpp"
#include "fwx_internal.hpp"

namespace fwx {

namespace {

constexpr int kPT = 64;        // P tile
constexpr int kPU = 32;        // u per chunk
constexpr int kPLA = kPU + 4;  // A row stride (dwords)

__device__ __forceinline__ uint32_t pred_key(uint32_t u, uint32_t t) {
  return __builtin_bit_cast(uint32_t, __builtin_amdgcn_cvt_pk_u16(u, t));
}

template <bool VEC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_predecessors(const int32_t* __restrict__ W, const int32_t* __restrict__ D,
                                                       int32_t* __restrict__ P, int32_t n) {
  __shared__ __align__(16) int32_t As[kPT * kPLA];  // D[i][u], row-major
  __shared__ __align__(16) int32_t Bs[kPU * kPT];   // W[u][j]
  const int tid = threadIdx.x;
  const int tx = tid & 15, ty = tid >> 4;
  const int i0 = blockIdx.y * kPT, j0 = blockIdx.x * kPT;
  uint32_t nd[4][4];
  uint32_t best[4][4];
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int i = i0 + ty + 16 * r, j = j0 + 4 * tx + c;
      nd[r][c] = (i < n && j < n) ? 0u - (uint32_t)D[(int64_t)i * n + j] : 0u;
      best[r][c] = 0xFFFFFFFFu;
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
#pragma unroll 1
    for (int u = 0; u < kPU; u += 4) {
      int4 a[4], b[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) a[r] = *reinterpret_cast<const int4*>(&As[(ty + 16 * r) * kPLA + u]);
#pragma unroll
      for (int e = 0; e < 4; ++e) b[e] = *reinterpret_cast<const int4*>(&Bs[(u + e) * kPT + 4 * tx]);
      const uint32_t uk = (uint32_t)(u0 + u);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const uint32_t ar[4] = {(uint32_t)a[r].x, (uint32_t)a[r].y, (uint32_t)a[r].z, (uint32_t)a[r].w};
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const uint32_t bc[4] = {(uint32_t)(&b[0].x)[c], (uint32_t)(&b[1].x)[c], (uint32_t)(&b[2].x)[c],
                                  (uint32_t)(&b[3].x)[c]};
          const uint32_t k0 = pred_key(uk + 0u, ar[0] + bc[0] + nd[r][c]);
          const uint32_t k1 = pred_key(uk + 1u, ar[1] + bc[1] + nd[r][c]);
          const uint32_t k2 = pred_key(uk + 2u, ar[2] + bc[2] + nd[r][c]);
          const uint32_t k3 = pred_key(uk + 3u, ar[3] + bc[3] + nd[r][c]);
          best[r][c] = vmin3u(vmin3u(best[r][c], k0, k1), k2, k3);
        }
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int i = i0 + ty + 16 * r;
    if (i >= n) continue;
    int32_t o[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int j = j0 + 4 * tx + c;
      o[c] = (i == j || (best[r][c] >> 16) != 0u) ? -1 : (int32_t)(best[r][c] & 0xFFFFu);
    }
    const int jb = j0 + 4 * tx;
    if (VEC && jb + 4 <= n) {
      *reinterpret_cast<int4*>(P + (int64_t)i * n + jb) = make_int4(o[0], o[1], o[2], o[3]);
    } else {
#pragma unroll
      for (int c = 0; c < 4; ++c)
        if (jb + c < n) P[(int64_t)i * n + jb + c] = o[c];
    }
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
  const bool vec = (n % 4 == 0) && ((reinterpret_cast<uintptr_t>(d_P) & 15u) == 0);
  if (vec)
    k_predecessors<true><<<dim3(nt, nt), 256, 0, st>>>(d_W, d_D, d_P, n);
  else
    k_predecessors<false><<<dim3(nt, nt), 256, 0, st>>>(d_W, d_D, d_P, n);
  FWX_LAUNCHED("k_predecessors");
  FWX_HIP(hipStreamSynchronize(st));
  return FWX_OK;
}

}  // extern "C"
