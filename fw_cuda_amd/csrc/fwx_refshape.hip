// fwx_refshape.hip -- three-phase blocked FW shaped like the reference's
// tiling_gpu kernels, kept as the like-for-like comparator on the same GPU.
//
// Shape of src/fw_tiling_gpu.cu:8-284: 32 x 32 tiles, 256 threads per block,
// each thread owning a 2 x 2 cell group (rows 2ty, 2ty+1; columns 2tx,
// 2tx+1), four launches per k-tile (diagonal, row, column, others), plain
// add + min arithmetic, the row and column phases as 32 dependent k steps,
// the others phase as an unrolled 32-deep independent update from LDS.
//
// The dependent phases publish row/column k through a double-buffered LDS
// line with one barrier per step (no in-place read/write overlap inside a
// step), so they stay exact even when a diagonal entry is negative.
#include <hip/hip_runtime.h>

#include "fwx_device.hpp"
#include "fwx_internal.hpp"

namespace fwx {

namespace {

constexpr int RT = 32;

__device__ __forceinline__ int32_t ld_or_inf(const int32_t* D, int32_t n, int32_t r, int32_t c) {
  return (r < n && c < n) ? D[(int64_t)r * n + c] : kInf;
}

// Diagonal tile: dependent closure.
__global__ __launch_bounds__(256) void k_ref_diag(int32_t* __restrict__ D, int32_t n, int32_t t) {
  __shared__ int32_t rowl[2][RT], coll[2][RT];
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const int32_t b = t * RT;
  int32_t c[2][2];
  for (int a = 0; a < 2; ++a)
    for (int e = 0; e < 2; ++e) c[a][e] = ld_or_inf(D, n, b + 2 * ty + a, b + 2 * tx + e);
  const int kmax = min(RT, n - b);
  for (int k = 0; k < kmax; ++k) {
    const int buf = k & 1;
    if (2 * ty == (k & ~1)) {
      rowl[buf][2 * tx] = c[k & 1][0];
      rowl[buf][2 * tx + 1] = c[k & 1][1];
    }
    if (2 * tx == (k & ~1)) {
      coll[buf][2 * ty] = c[0][k & 1];
      coll[buf][2 * ty + 1] = c[1][k & 1];
    }
    __syncthreads();
    for (int a = 0; a < 2; ++a)
      for (int e = 0; e < 2; ++e)
        c[a][e] = min(c[a][e], wadd(coll[buf][2 * ty + a], rowl[buf][2 * tx + e]));
  }
  for (int a = 0; a < 2; ++a)
    for (int e = 0; e < 2; ++e)
      if (b + 2 * ty + a < n && b + 2 * tx + e < n)
        D[(int64_t)(b + 2 * ty + a) * n + b + 2 * tx + e] = c[a][e];
}

// Row (ROW=true: tile (t, j)) or column (tile (i, t)) phase: 32 dependent
// steps against the closed diagonal tile held in LDS.
template <bool ROW>
__global__ __launch_bounds__(256) void k_ref_rowcol(int32_t* __restrict__ D, int32_t n, int32_t t) {
  __shared__ int32_t dg[RT][RT];
  __shared__ int32_t line[2][RT];
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const int32_t db = t * RT;
  int other = blockIdx.x;
  other += (other >= t);
  const int32_t r0 = ROW ? db : other * RT;
  const int32_t c0 = ROW ? other * RT : db;
  for (int a = 0; a < 2; ++a)
    for (int e = 0; e < 2; ++e)
      dg[2 * ty + a][2 * tx + e] = ld_or_inf(D, n, db + 2 * ty + a, db + 2 * tx + e);
  int32_t c[2][2];
  for (int a = 0; a < 2; ++a)
    for (int e = 0; e < 2; ++e) c[a][e] = ld_or_inf(D, n, r0 + 2 * ty + a, c0 + 2 * tx + e);
  __syncthreads();
  const int kmax = min(RT, n - db);
  for (int k = 0; k < kmax; ++k) {
    const int buf = k & 1;
    if (ROW) {  // C[i][j] = min(C[i][j], Dg[i][k] + C[k][j]); publish C row k
      if (2 * ty == (k & ~1)) {
        line[buf][2 * tx] = c[k & 1][0];
        line[buf][2 * tx + 1] = c[k & 1][1];
      }
    } else {  // C[i][j] = min(C[i][j], C[i][k] + Dg[k][j]); publish C column k
      if (2 * tx == (k & ~1)) {
        line[buf][2 * ty] = c[0][k & 1];
        line[buf][2 * ty + 1] = c[1][k & 1];
      }
    }
    __syncthreads();
    for (int a = 0; a < 2; ++a)
      for (int e = 0; e < 2; ++e) {
        const int32_t v = ROW ? wadd(dg[2 * ty + a][k], line[buf][2 * tx + e])
                              : wadd(line[buf][2 * ty + a], dg[k][2 * tx + e]);
        c[a][e] = min(c[a][e], v);
      }
  }
  for (int a = 0; a < 2; ++a)
    for (int e = 0; e < 2; ++e)
      if (r0 + 2 * ty + a < n && c0 + 2 * tx + e < n)
        D[(int64_t)(r0 + 2 * ty + a) * n + c0 + 2 * tx + e] = c[a][e];
}

// Others phase: tile (i, j), i, j != t: independent 32-deep update.
__global__ __launch_bounds__(256) void k_ref_others(int32_t* __restrict__ D, int32_t n, int32_t t,
                                                    int32_t N) {
  __shared__ int32_t sa[RT][RT], sb[RT][RT];
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  int bi = blockIdx.y, bj = blockIdx.x;
  bi += (bi >= t);
  bj += (bj >= t);
  const int32_t db = t * RT, r0 = bi * RT, c0 = bj * RT;
  for (int a = 0; a < 2; ++a)
    for (int e = 0; e < 2; ++e) {
      sa[2 * ty + a][2 * tx + e] = ld_or_inf(D, n, r0 + 2 * ty + a, db + 2 * tx + e);
      sb[2 * ty + a][2 * tx + e] = ld_or_inf(D, n, db + 2 * ty + a, c0 + 2 * tx + e);
    }
  int32_t c[2][2];
  for (int a = 0; a < 2; ++a)
    for (int e = 0; e < 2; ++e) c[a][e] = ld_or_inf(D, n, r0 + 2 * ty + a, c0 + 2 * tx + e);
  __syncthreads();
#pragma unroll
  for (int k = 0; k < RT; ++k)
    for (int a = 0; a < 2; ++a)
      for (int e = 0; e < 2; ++e) c[a][e] = min(c[a][e], wadd(sa[2 * ty + a][k], sb[k][2 * tx + e]));
  for (int a = 0; a < 2; ++a)
    for (int e = 0; e < 2; ++e)
      if (r0 + 2 * ty + a < n && c0 + 2 * tx + e < n)
        D[(int64_t)(r0 + 2 * ty + a) * n + c0 + 2 * tx + e] = c[a][e];
}

}  // namespace

int enqueue_refshape(int32_t* d, int32_t n, hipStream_t st) {
  const int N = (n + RT - 1) / RT;
  for (int t = 0; t < N; ++t) {
    k_ref_diag<<<1, 256, 0, st>>>(d, n, t);
    FWX_LAUNCHED("k_ref_diag");
    if (N > 1) {
      k_ref_rowcol<true><<<N - 1, 256, 0, st>>>(d, n, t);
      FWX_LAUNCHED("k_ref_rowcol<row>");
      k_ref_rowcol<false><<<N - 1, 256, 0, st>>>(d, n, t);
      FWX_LAUNCHED("k_ref_rowcol<col>");
      k_ref_others<<<dim3(N - 1, N - 1), 256, 0, st>>>(d, n, t, N);
      FWX_LAUNCHED("k_ref_others");
    }
  }
  return FWX_OK;
}

}  // namespace fwx
