// fwx_diff.hip -- differentiable fp32 Floyd-Warshall (no reference counterpart).
//
// Forward: D = min-plus closure of W (diagonal taken as 0, +inf = no edge),
// blocked three-phase FW over 64 x 64 tiles, batched over graphs (grid.y),
// carrying the predecessor P[i][j] (vertex before j on the chosen i->j path)
// through every product: when A[i][k] + B[k][j] < C[i][j] strictly, the
// cell takes B's predecessor of (k, j).  Equal candidates keep the earlier
// one, so ties resolve deterministically; with continuous weights they have
// probability ~0.
//
// Backward of L = sum(G * D): each D[s][t] is the sum of the weights on the
// chosen s->t path, so dL/dW[u][v] = sum of G[s][t] over the (s,t) whose path
// uses edge (u,v).  Per source s (one workgroup) the predecessor row is a
// tree rooted at s; a vertex's accumulated gradient g[t] = G[s][t] + sum of
// its children's g is added to dW[P[s][t]][t].  The tree is processed level
// by level from the deepest (depths by pointer jumping), so all children of
// a vertex are final before it is pushed to its parent.
//
// Without a stored P, the backward recomputes each source block's
// predecessors as P[s][t] = argmin_u D[s][u] + W[u][t] (u != t): the last
// edge of a shortest path satisfies D[s][t] = D[s][u] + W[u][t].  It needs a
// 256 x n int32 workspace instead of n x n.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <mutex>

#include "fwx_internal.hpp"

namespace fwx {

namespace {

constexpr int FT = 64;   // tile
constexpr int FKC = 32;  // k chunk
constexpr int FLDA = FKC + 4;
// Chunk depth of the fast rest product.  16 halves its LDS (25.6 KB per
// block, up to 5 blocks per CU instead of 3) for twice the barriers; at
// n=4096 the forward takes 8.55 ms with 16 and 9.40 ms with 32.
#ifndef FWX_DIFF_KC
#define FWX_DIFF_KC 16
#endif
constexpr int kRestKC = FWX_DIFF_KC;
constexpr float kFInf = __builtin_huge_valf();

// Window of a row-major fp32 matrix and its predecessor matrix.
struct FView {
  float* p;
  int32_t* pp;  // may be null
  int32_t ld, rlim, clim;
  __device__ __forceinline__ bool in(int r, int c) const { return r < rlim && c < clim; }
};

__device__ __forceinline__ FView fview(float* D, int32_t* P, int32_t n, int32_t r0, int32_t c0) {
  const int64_t o = (int64_t)r0 * n + c0;
  return FView{D + o, P ? P + o : nullptr, n, n - r0, n - c0};
}

struct FSmem {
  float a[FT * FLDA];
  float b[FKC * FT];
  int32_t bp[FKC * FT];
};

// acc <- min(acc, A (x) B) over depth k values, predecessor from B's
// predecessors (PRED) or, when BP_IS_ROW, the operand row index k itself.
// ExcludeDiag: B[k][j] with k == j (in global indices) is treated as +inf.
template <bool PRED, bool BP_IS_ROW>
__device__ void fprod(FSmem& s, const FView& A, const FView& B, int32_t depth, int32_t brow0, int32_t bcol0,
                      bool exclude_diag, float (&acc)[4][4], int32_t (&pacc)[4][4]) {
  const int tid = threadIdx.x;
  const int tx = tid & 15, ty = tid >> 4;
  for (int kc = 0; kc < depth; kc += FKC) {
    // stage A: 64 rows x 32 k -> 8 per thread
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int idx = e * 256 + tid;
      const int r = idx / FKC, k = idx % FKC;
      s.a[r * FLDA + k] = A.in(r, kc + k) ? A.p[r * A.ld + kc + k] : kFInf;
    }
    // stage B: 32 k x 64 cols -> 8 per thread
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int idx = e * 256 + tid;
      const int k = idx / FT, c = idx % FT;
      const bool ok = B.in(kc + k, c) && !(exclude_diag && brow0 + kc + k == bcol0 + c);
      s.b[k * FT + c] = ok ? B.p[(kc + k) * B.ld + c] : kFInf;
      if (PRED) {
        if (BP_IS_ROW)
          s.bp[k * FT + c] = brow0 + kc + k;
        else
          s.bp[k * FT + c] = ok ? B.pp[(kc + k) * B.ld + c] : -1;
      }
    }
    __syncthreads();
#pragma unroll 4
    for (int k = 0; k < FKC; ++k) {
      float a[4], b[4];
      int32_t bp[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) a[r] = s.a[(ty + 16 * r) * FLDA + k];
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        b[c] = s.b[k * FT + tx + 16 * c];
        if (PRED) bp[c] = s.bp[k * FT + tx + 16 * c];
      }
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const float cand = a[r] + b[c];
          if (cand < acc[r][c]) {
            acc[r][c] = cand;
            if (PRED) pacc[r][c] = bp[c];
          }
        }
    }
    __syncthreads();
  }
}

template <bool PRED>
__device__ __forceinline__ void fload(const FView& C, float (&acc)[4][4], int32_t (&pacc)[4][4]) {
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int rr = ty + 16 * r, cc = tx + 16 * c;
      const bool ok = C.in(rr, cc);
      acc[r][c] = ok ? C.p[rr * C.ld + cc] : kFInf;
      if (PRED) pacc[r][c] = ok ? C.pp[rr * C.ld + cc] : -1;
    }
}

template <bool PRED>
__device__ __forceinline__ void fstore(const FView& C, const float (&acc)[4][4], const int32_t (&pacc)[4][4]) {
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int rr = ty + 16 * r, cc = tx + 16 * c;
      if (C.in(rr, cc)) {
        C.p[rr * C.ld + cc] = acc[r][c];
        if (PRED) C.pp[rr * C.ld + cc] = pacc[r][c];
      }
    }
}

// D = W with a zero diagonal; P = i for a present edge, else -1.
template <bool PRED>
__global__ __launch_bounds__(256) void k_finit(const float* __restrict__ W, float* __restrict__ D,
                                               int32_t* __restrict__ P, int32_t n, int64_t total) {
  const int64_t nn = (int64_t)n * n;
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < total; e += (int64_t)gridDim.x * 256) {
    const int64_t c = e % nn;
    const int32_t i = (int32_t)(c / n), j = (int32_t)(c % n);
    const float w = W[e];
    const bool edge = i != j && w < kFInf;
    D[e] = i == j ? 0.0f : w;
    if (PRED) P[e] = edge ? i : -1;
  }
}

template <bool PRED>
__global__ __launch_bounds__(256) void k_fdiag(float* __restrict__ D, int32_t* __restrict__ P, int32_t n,
                                               int32_t c) {
  __shared__ float rowv[2][FT], colv[2][FT];
  __shared__ int32_t rowp[2][FT];
  const int64_t g = blockIdx.y;
  float* Dg = D + g * (int64_t)n * n;
  int32_t* Pg = PRED ? P + g * (int64_t)n * n : nullptr;
  const FView V = fview(Dg, Pg, n, c * FT, c * FT);
  float acc[4][4];
  int32_t pacc[4][4];
  fload<PRED>(V, acc, pacc);
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const int kmax = min(FT, n - c * FT);
#pragma unroll
  for (int kk = 0; kk < 4; ++kk) {
    for (int k16 = 0; k16 < 16; ++k16) {
      const int k = kk * 16 + k16;
      if (k >= kmax) break;
      const int b = k & 1;
      if (ty == k16) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          rowv[b][tx + 16 * q] = acc[kk][q];
          if (PRED) rowp[b][tx + 16 * q] = pacc[kk][q];
        }
      }
      if (tx == k16) {
#pragma unroll
        for (int r = 0; r < 4; ++r) colv[b][ty + 16 * r] = acc[r][kk];
      }
      __syncthreads();
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const float ck = colv[b][ty + 16 * r];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const float cand = ck + rowv[b][tx + 16 * q];
          if (cand < acc[r][q]) {
            acc[r][q] = cand;
            if (PRED) pacc[r][q] = rowp[b][tx + 16 * q];
          }
        }
      }
    }
  }
  fstore<PRED>(V, acc, pacc);
}

// Diagonal closure, 1024 threads (one row and 4 columns each) and two k
// steps per barrier.  The owners of rows k, k+1 and columns k, k+1 publish
// them (values, and the rows' predecessors); every thread then forms column
// k+1 after step k for its row and row k+1 after step k for its columns --
// the same candidates, sums and strict comparisons as the sequential loop --
// and applies step k, then step k+1.  Same bits as k_fdiag
// (FWX_DIFF_DIAG=0 selects it).  Line buffer p&1 is rewritten at pair p+2,
// after every thread passed barrier p+1.
template <bool PRED>
__global__ __launch_bounds__(1024) void k_fdiag2(float* __restrict__ D, int32_t* __restrict__ P, int32_t n,
                                                 int32_t c, int32_t prio) {
  if (prio) __builtin_amdgcn_s_setprio(3);
  __shared__ float rowv[2][2][FT], colv[2][2][FT];
  __shared__ int32_t rowp[2][2][FT];
  const int64_t g = blockIdx.y;
  float* Dg = D + g * (int64_t)n * n;
  int32_t* Pg = PRED ? P + g * (int64_t)n * n : nullptr;
  const FView V = fview(Dg, Pg, n, c * FT, c * FT);
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  float acc[4];
  int32_t pacc[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const bool in = V.in(ty, tx + 16 * q);
    acc[q] = in ? V.p[ty * V.ld + tx + 16 * q] : kFInf;
    if (PRED) pacc[q] = in ? V.pp[ty * V.ld + tx + 16 * q] : -1;
  }
  const int kmax = min(FT, n - c * FT);
#pragma unroll
  for (int qk = 0; qk < 4; ++qk) {
    for (int k16 = 0; k16 < 16; k16 += 2) {
      const int k = qk * 16 + k16;
      if (k >= kmax) break;
      const int b = (k >> 1) & 1;
      if (ty == k || ty == k + 1) {
        const int h = ty - k;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          rowv[b][h][tx + 16 * q] = acc[q];
          if (PRED) rowp[b][h][tx + 16 * q] = pacc[q];
        }
      }
      if (tx == k16 || tx == k16 + 1) colv[b][tx - k16][ty] = acc[qk];
      __syncthreads();
      const float ci0 = colv[b][0][ty];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const float cand = ci0 + rowv[b][0][tx + 16 * q];
        if (cand < acc[q]) {
          acc[q] = cand;
          if (PRED) pacc[q] = rowp[b][0][tx + 16 * q];
        }
      }
      if (k + 1 < kmax) {
        const float dkk1 = rowv[b][0][k + 1];  // D[k][k+1]
        const float dk1k = rowv[b][1][k];      // D[k+1][k]
        float ci1 = colv[b][1][ty];             // D[ty][k+1] after step k
        const float t = ci0 + dkk1;
        if (t < ci1) ci1 = t;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int j = tx + 16 * q;
          float r1 = rowv[b][1][j];  // D[k+1][j] after step k
          int32_t p1 = PRED ? rowp[b][1][j] : 0;
          const float u = dk1k + rowv[b][0][j];
          if (u < r1) {
            r1 = u;
            if (PRED) p1 = rowp[b][0][j];
          }
          const float cand = ci1 + r1;
          if (cand < acc[q]) {
            acc[q] = cand;
            if (PRED) pacc[q] = p1;
          }
        }
      }
    }
  }
#pragma unroll
  for (int q = 0; q < 4; ++q)
    if (V.in(ty, tx + 16 * q)) {
      V.p[ty * V.ld + tx + 16 * q] = acc[q];
      if (PRED) V.pp[ty * V.ld + tx + 16 * q] = pacc[q];
    }
}

template <bool PRED>
__global__ __launch_bounds__(256) void k_fpanel(float* __restrict__ D, int32_t* __restrict__ P, int32_t n,
                                                int32_t c, int32_t N, int32_t prio) {
  if (prio) __builtin_amdgcn_s_setprio(3);
  __shared__ FSmem s;
  const int64_t g = blockIdx.y;
  float* Dg = D + g * (int64_t)n * n;
  int32_t* Pg = PRED ? P + g * (int64_t)n * n : nullptr;
  const int b = blockIdx.x;
  const FView diag = fview(Dg, Pg, n, c * FT, c * FT);
  FView A, B, C;
  if (b < N - 1) {
    const int j = b + (b >= c);
    C = fview(Dg, Pg, n, c * FT, j * FT);
    A = diag;
    B = C;
  } else {
    int i = b - (N - 1);
    i += (i >= c);
    C = fview(Dg, Pg, n, i * FT, c * FT);
    A = C;
    B = diag;
  }
  float acc[4][4];
  int32_t pacc[4][4];
  fload<PRED>(C, acc, pacc);
  fprod<PRED, false>(s, A, B, FT, 0, 0, false, acc, pacc);
  fstore<PRED>(C, acc, pacc);
}

// Rest tile (i, j) of block b in stage c.  mode 0: every (i, j) with
// i, j != c, (N-1)^2 blocks.  mode 1: the band stage c+1's closure and
// panels read -- row c+1 (j != c, N-1 blocks) and column c+1 (i != c, c+1,
// N-2 blocks).  mode 2: the others (i, j not in {c, c+1}), (N-2)^2 blocks.
__device__ __forceinline__ void rest_tile(int b, int c, int N, int mode, int* I, int* J) {
  int i, j;
  if (mode == 0) {
    i = b / (N - 1);
    j = b % (N - 1);
    i += (i >= c);
    j += (j >= c);
  } else if (mode == 1) {
    if (b < N - 1) {
      i = c + 1;
      j = b + (b >= c);
    } else {
      i = b - (N - 1);
      i = i < c ? i : i + 2;
      j = c + 1;
    }
  } else {
    i = b / (N - 2);
    j = b % (N - 2);
    i = i < c ? i : i + 2;
    j = j < c ? j : j + 2;
  }
  *I = i;
  *J = j;
}

template <bool PRED>
__global__ __launch_bounds__(256) void k_frest(float* __restrict__ D, int32_t* __restrict__ P, int32_t n,
                                               int32_t c, int32_t N, int32_t mode) {
  __shared__ FSmem s;
  const int64_t g = blockIdx.y;
  float* Dg = D + g * (int64_t)n * n;
  int32_t* Pg = PRED ? P + g * (int64_t)n * n : nullptr;
  int i, j;
  rest_tile(blockIdx.x, c, N, mode, &i, &j);
  const FView A = fview(Dg, Pg, n, i * FT, c * FT);
  const FView B = fview(Dg, Pg, n, c * FT, j * FT);
  const FView C = fview(Dg, Pg, n, i * FT, j * FT);
  float acc[4][4];
  int32_t pacc[4][4];
  fload<PRED>(C, acc, pacc);
  fprod<PRED, false>(s, A, B, FT, 0, 0, false, acc, pacc);
  fstore<PRED>(C, acc, pacc);
}

// ---------------------------------------------------------------------------
// Fast rest-tile product (k_frest4).  Same arithmetic and candidate order as
// fprod -- per cell, k ascending, a + b in fp32, strict < -- so D and P are
// bit-identical; laid out for 16-byte LDS reads.  Thread (tx, ty) owns rows
// ty + 16 r and the 4 adjacent columns 4 tx + c; per 4 k it reads 4 A rows
// (one float4 each: 4 k values), and 4 B and 4 predecessor rows (one
// float4 / int4 each: its 4 columns).  The chunks (kRestKC deep) are
// double-buffered: the next chunk's global loads are in flight while the
// current one is computed.
template <int KC>
struct alignas(16) FSmem2 {
  alignas(16) float a[2][FT * (KC + 4)];
  alignas(16) float b[2][KC * FT];
  alignas(16) int32_t bp[2][KC * FT];
};

template <bool PRED, bool VEC, int KC>
struct FProd4 {
  static constexpr int LDA = KC + 4;
  static constexpr int AQ = KC / 4;         // 4-k quads per A row
  static constexpr int NP = KC / 16;        // staging units per thread (A and B alike)
  static_assert(KC % 16 == 0, "chunk of 16 or 32");
  // staging: A 64 rows x AQ quads, B / BP KC rows x 16 quads
  float4 ra[NP], rb[NP];
  int4 rp[NP];

  __device__ __forceinline__ void fetch(const FView& A, const FView& B, int32_t kc) {
    const int tid = threadIdx.x;
    const float4 inf4 = make_float4(kFInf, kFInf, kFInf, kFInf);
#pragma unroll
    for (int p = 0; p < NP; ++p) {
      const int u = p * 256 + tid;
      const int r = u / AQ, k = kc + (u % AQ) * 4;
      if (VEC) {
        ra[p] = (r < A.rlim && k < A.clim) ? *reinterpret_cast<const float4*>(A.p + r * A.ld + k) : inf4;
      } else {
        float v[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) v[e] = (r < A.rlim && k + e < A.clim) ? A.p[r * A.ld + k + e] : kFInf;
        ra[p] = make_float4(v[0], v[1], v[2], v[3]);
      }
      const int kr = kc + (u >> 4), col = (u & 15) * 4;
      if (VEC) {
        const bool in = kr < B.rlim && col < B.clim;
        rb[p] = in ? *reinterpret_cast<const float4*>(B.p + kr * B.ld + col) : inf4;
        if (PRED) rp[p] = in ? *reinterpret_cast<const int4*>(B.pp + kr * B.ld + col) : make_int4(-1, -1, -1, -1);
      } else {
        float v[4];
        int32_t q[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const bool in = kr < B.rlim && col + e < B.clim;
          v[e] = in ? B.p[kr * B.ld + col + e] : kFInf;
          if (PRED) q[e] = in ? B.pp[kr * B.ld + col + e] : -1;
        }
        rb[p] = make_float4(v[0], v[1], v[2], v[3]);
        if (PRED) rp[p] = make_int4(q[0], q[1], q[2], q[3]);
      }
    }
  }

  __device__ __forceinline__ void stash(FSmem2<KC>& s, int buf) const {
    const int tid = threadIdx.x;
#pragma unroll
    for (int p = 0; p < NP; ++p) {
      const int u = p * 256 + tid;
      *reinterpret_cast<float4*>(&s.a[buf][(u / AQ) * LDA + (u % AQ) * 4]) = ra[p];
      *reinterpret_cast<float4*>(&s.b[buf][(u >> 4) * FT + (u & 15) * 4]) = rb[p];
      if (PRED) *reinterpret_cast<int4*>(&s.bp[buf][(u >> 4) * FT + (u & 15) * 4]) = rp[p];
    }
  }

  __device__ __forceinline__ static void compute(const FSmem2<KC>& s, int buf, float (&acc)[4][4],
                                                 int32_t (&pacc)[4][4]) {
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
#pragma unroll 2
    for (int q = 0; q < KC / 4; ++q) {
      float4 a4[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) a4[r] = *reinterpret_cast<const float4*>(&s.a[buf][(ty + 16 * r) * LDA + 4 * q]);
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const float4 b4 = *reinterpret_cast<const float4*>(&s.b[buf][(4 * q + e) * FT + 4 * tx]);
        int4 p4;
        if (PRED) p4 = *reinterpret_cast<const int4*>(&s.bp[buf][(4 * q + e) * FT + 4 * tx]);
        const float bv[4] = {b4.x, b4.y, b4.z, b4.w};
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const float ae = (&a4[r].x)[e];
#pragma unroll
          for (int c = 0; c < 4; ++c) {
            // The value is min(acc, cand) (v_min_f32: no NaN, and a
            // candidate equal to acc leaves the same value), so only the
            // predecessor select reads the strict-< mask, and the min can
            // fill the wait states the mask needs.
            const float cand = ae + bv[c];
            if (PRED && cand < acc[r][c]) pacc[r][c] = (&p4.x)[c];
            acc[r][c] = fminf(acc[r][c], cand);
          }
        }
      }
    }
  }
};

template <bool PRED, bool VEC>
__device__ __forceinline__ void fload4(const FView& C, float (&acc)[4][4], int32_t (&pacc)[4][4]) {
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int rr = ty + 16 * r, cc = 4 * tx;
    if (VEC) {
      const bool in = rr < C.rlim && cc < C.clim;
      const float4 v = in ? *reinterpret_cast<const float4*>(C.p + rr * C.ld + cc)
                          : make_float4(kFInf, kFInf, kFInf, kFInf);
      acc[r][0] = v.x, acc[r][1] = v.y, acc[r][2] = v.z, acc[r][3] = v.w;
      if (PRED) {
        const int4 q = in ? *reinterpret_cast<const int4*>(C.pp + rr * C.ld + cc) : make_int4(-1, -1, -1, -1);
        pacc[r][0] = q.x, pacc[r][1] = q.y, pacc[r][2] = q.z, pacc[r][3] = q.w;
      }
    } else {
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const bool in = C.in(rr, cc + c);
        acc[r][c] = in ? C.p[rr * C.ld + cc + c] : kFInf;
        if (PRED) pacc[r][c] = in ? C.pp[rr * C.ld + cc + c] : -1;
      }
    }
  }
}

template <bool PRED, bool VEC>
__device__ __forceinline__ void fstore4(const FView& C, const float (&acc)[4][4], const int32_t (&pacc)[4][4]) {
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int rr = ty + 16 * r, cc = 4 * tx;
    if (VEC) {
      if (rr < C.rlim && cc < C.clim) {
        *reinterpret_cast<float4*>(C.p + rr * C.ld + cc) = make_float4(acc[r][0], acc[r][1], acc[r][2], acc[r][3]);
        if (PRED)
          *reinterpret_cast<int4*>(C.pp + rr * C.ld + cc) = make_int4(pacc[r][0], pacc[r][1], pacc[r][2], pacc[r][3]);
      }
    } else {
#pragma unroll
      for (int c = 0; c < 4; ++c)
        if (C.in(rr, cc + c)) {
          C.p[rr * C.ld + cc + c] = acc[r][c];
          if (PRED) C.pp[rr * C.ld + cc + c] = pacc[r][c];
        }
    }
  }
}

// Rest tiles of stage c with FProd4 (same tile decode as k_frest).
template <bool PRED, bool VEC>
__global__ __launch_bounds__(256) void k_frest4(float* __restrict__ D, int32_t* __restrict__ P, int32_t n,
                                                int32_t c, int32_t N, int32_t mode, int32_t prio) {
  if (prio) __builtin_amdgcn_s_setprio(3);
  __shared__ FSmem2<kRestKC> s;
  const int64_t g = blockIdx.y;
  float* Dg = D + g * (int64_t)n * n;
  int32_t* Pg = PRED ? P + g * (int64_t)n * n : nullptr;
  int i, j;
  rest_tile(blockIdx.x, c, N, mode, &i, &j);
  const FView A = fview(Dg, Pg, n, i * FT, c * FT);
  const FView B = fview(Dg, Pg, n, c * FT, j * FT);
  const FView C = fview(Dg, Pg, n, i * FT, j * FT);
  float acc[4][4];
  int32_t pacc[4][4];
  fload4<PRED, VEC>(C, acc, pacc);
  FProd4<PRED, VEC, kRestKC> pr;
  pr.fetch(A, B, 0);
  pr.stash(s, 0);
  __syncthreads();
  constexpr int nch = FT / kRestKC;
#pragma unroll
  for (int ch = 0; ch < nch; ++ch) {
    if (ch + 1 < nch) pr.fetch(A, B, (ch + 1) * kRestKC);
    FProd4<PRED, VEC, kRestKC>::compute(s, ch & 1, acc, pacc);
    if (ch + 1 < nch) pr.stash(s, (ch + 1) & 1);
    __syncthreads();
  }
  fstore4<PRED, VEC>(C, acc, pacc);
}

// Panels of stage c with FProd4 (same tiles and operands as k_fpanel: the
// row/column tile's own pre-update values are one operand, staged chunk by
// chunk before the block stores anything).
template <bool PRED, bool VEC>
__global__ __launch_bounds__(256) void k_fpanel4(float* __restrict__ D, int32_t* __restrict__ P, int32_t n,
                                                 int32_t c, int32_t N, int32_t prio) {
  if (prio) __builtin_amdgcn_s_setprio(3);
  __shared__ FSmem2<kRestKC> s;
  const int64_t g = blockIdx.y;
  float* Dg = D + g * (int64_t)n * n;
  int32_t* Pg = PRED ? P + g * (int64_t)n * n : nullptr;
  const int b = blockIdx.x;
  const FView diag = fview(Dg, Pg, n, c * FT, c * FT);
  FView A, B, C;
  if (b < N - 1) {
    const int j = b + (b >= c);
    C = fview(Dg, Pg, n, c * FT, j * FT);
    A = diag;
    B = C;
  } else {
    int i = b - (N - 1);
    i += (i >= c);
    C = fview(Dg, Pg, n, i * FT, c * FT);
    A = C;
    B = diag;
  }
  float acc[4][4];
  int32_t pacc[4][4];
  fload4<PRED, VEC>(C, acc, pacc);
  FProd4<PRED, VEC, kRestKC> pr;
  pr.fetch(A, B, 0);
  pr.stash(s, 0);
  __syncthreads();
  constexpr int nch = FT / kRestKC;
#pragma unroll
  for (int ch = 0; ch < nch; ++ch) {
    if (ch + 1 < nch) pr.fetch(A, B, (ch + 1) * kRestKC);
    FProd4<PRED, VEC, kRestKC>::compute(s, ch & 1, acc, pacc);
    if (ch + 1 < nch) pr.stash(s, (ch + 1) & 1);
    __syncthreads();
  }
  fstore4<PRED, VEC>(C, acc, pacc);
}

// ---------------------------------------------------------------------------
// Multi-layer stages (kappa k-tiles per stage, FWX_DIFF_KAPPA).  Per stage
// of k-tiles [e0, e0+en): for each c in it, the closure of (c, c), the
// panels of row and column c over every tile, and the cross update of the
// stage's other row/column tiles with k-tile c; then one rest pass of depth
// en*64 over every tile outside the stage's rows and columns.  Each product
// walks its k ascending with the same strict < rule, so ties keep the
// earlier candidate as before; sums may round differently from the
// one-k-tile schedule (a different decomposition of the same paths).

// Tile (i, j) of block b of a multi-layer rest pass.  Stage tiles E =
// [e0, e0+en), next stage F = [f0, f0+fn) with f0 = e0 + en.
//   mode 0: i, j not in E, (N-en)^2 blocks;
//   mode 1: the band F's lead reads -- i in F, j not in E (fn (N-en)
//           blocks), then i not in E or F, j in F ((N-en-fn) fn blocks);
//   mode 2: i, j not in E or F, (N-en-fn)^2 blocks.
__device__ __forceinline__ void rest_tile_ml(int b, int e0, int en, int fn, int N, int mode, int* I, int* J) {
  int i, j;
  if (mode == 0) {
    const int m = N - en;
    i = b / m;
    j = b % m;
    i += (i >= e0) ? en : 0;
    j += (j >= e0) ? en : 0;
  } else if (mode == 1) {
    const int m = N - en;
    const int f0 = e0 + en;
    if (b < fn * m) {
      i = f0 + b / m;
      j = b % m;
      j += (j >= e0) ? en : 0;
    } else {
      b -= fn * m;
      i = b / fn;
      j = f0 + b % fn;
      i += (i >= e0) ? en + fn : 0;
    }
  } else {
    const int m = N - en - fn;
    i = b / m;
    j = b % m;
    i += (i >= e0) ? en + fn : 0;
    j += (j >= e0) ? en + fn : 0;
  }
  *I = i;
  *J = j;
}

template <bool PRED, bool VEC>
__device__ __forceinline__ void ftile_product(FSmem2<kRestKC>& s, const FView& A, const FView& B, const FView& C,
                                              int32_t depth) {
  float acc[4][4];
  int32_t pacc[4][4];
  fload4<PRED, VEC>(C, acc, pacc);
  FProd4<PRED, VEC, kRestKC> pr;
  pr.fetch(A, B, 0);
  pr.stash(s, 0);
  __syncthreads();
  const int nch = depth / kRestKC;
  for (int ch = 0; ch < nch; ++ch) {
    if (ch + 1 < nch) pr.fetch(A, B, (ch + 1) * kRestKC);
    FProd4<PRED, VEC, kRestKC>::compute(s, ch & 1, acc, pacc);
    if (ch + 1 < nch) pr.stash(s, (ch + 1) & 1);
    __syncthreads();
  }
  fstore4<PRED, VEC>(C, acc, pacc);
}

template <bool PRED, bool VEC>
__global__ __launch_bounds__(256) void k_frest_ml(float* __restrict__ D, int32_t* __restrict__ P, int32_t n,
                                                  int32_t e0, int32_t en, int32_t fn, int32_t N, int32_t mode,
                                                  int32_t prio) {
  if (prio) __builtin_amdgcn_s_setprio(3);
  __shared__ FSmem2<kRestKC> s;
  const int64_t g = blockIdx.y;
  float* Dg = D + g * (int64_t)n * n;
  int32_t* Pg = PRED ? P + g * (int64_t)n * n : nullptr;
  int i, j;
  rest_tile_ml(blockIdx.x, e0, en, fn, N, mode, &i, &j);
  ftile_product<PRED, VEC>(s, fview(Dg, Pg, n, i * FT, e0 * FT), fview(Dg, Pg, n, e0 * FT, j * FT),
                           fview(Dg, Pg, n, i * FT, j * FT), en * FT);
}

// Cross update of stage [kb, kb+Bst) with k-tile c: stage rows (except c) x
// every column (except c), (Bst-1)(N-1) tiles, then other rows x stage
// columns (except c), (N-Bst)(Bst-1) tiles.
template <bool PRED, bool VEC>
__global__ __launch_bounds__(256) void k_fcross(float* __restrict__ D, int32_t* __restrict__ P, int32_t n,
                                                int32_t c, int32_t kb, int32_t Bst, int32_t N, int32_t prio) {
  if (prio) __builtin_amdgcn_s_setprio(3);
  __shared__ FSmem2<kRestKC> s;
  const int64_t g = blockIdx.y;
  float* Dg = D + g * (int64_t)n * n;
  int32_t* Pg = PRED ? P + g * (int64_t)n * n : nullptr;
  int b = blockIdx.x;
  const int l = c - kb;
  const int r1 = (Bst - 1) * (N - 1);
  int i, j;
  if (b < r1) {
    int m = b / (N - 1);
    const int jj = b % (N - 1);
    m += (m >= l);
    i = kb + m;
    j = jj + (jj >= c);
  } else {
    b -= r1;
    const int ii = b / (Bst - 1);
    int m = b % (Bst - 1);
    m += (m >= l);
    j = kb + m;
    i = ii + (ii >= kb ? Bst : 0);
  }
  ftile_product<PRED, VEC>(s, fview(Dg, Pg, n, i * FT, c * FT), fview(Dg, Pg, n, c * FT, j * FT),
                           fview(Dg, Pg, n, i * FT, j * FT), FT);
}

// Predecessors of sources [s0, s0+rows) recomputed from D and W:
// Pw[r][t] = argmin_{u != t} D[s0+r][u] + W[u][t], -1 for t == s or
// unreachable t.  Grid: (ceil(rows/64) * ceil(n/64), batch).
__global__ __launch_bounds__(256) void k_fpred(const float* __restrict__ W, const float* __restrict__ D,
                                               int32_t* __restrict__ Pw, int32_t n, int32_t s0, int32_t rows) {
  __shared__ FSmem s;
  const int64_t g = blockIdx.y;
  const int nt = (n + FT - 1) / FT;
  const int bi = blockIdx.x / nt, bj = blockIdx.x % nt;
  float* Dg = const_cast<float*>(D) + g * (int64_t)n * n;
  float* Wg = const_cast<float*>(W) + g * (int64_t)n * n;
  FView A = fview(Dg, nullptr, n, s0 + bi * FT, 0);
  A.rlim = min(A.rlim, rows - bi * FT);
  const FView B = fview(Wg, nullptr, n, 0, bj * FT);
  float acc[4][4];
  int32_t pacc[4][4];
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      acc[r][c] = kFInf;
      pacc[r][c] = -1;
    }
  fprod<true, true>(s, A, B, n + (FKC - n % FKC) % FKC, 0, bj * FT, true, acc, pacc);
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  int32_t* Pg = Pw + g * (int64_t)rows * n;
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int lr = bi * FT + ty + 16 * r, t = bj * FT + tx + 16 * c;
      if (lr < rows && t < n) {
        const int32_t src = s0 + lr;
        Pg[(int64_t)lr * n + t] = (t == src || !(acc[r][c] < kFInf)) ? -1 : pacc[r][c];
      }
    }
}

__global__ __launch_bounds__(256) void k_ffill(float* __restrict__ x, int64_t total) {
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < total; e += (int64_t)gridDim.x * 256) x[e] = 0.0f;
}

// Backward for sources [s0, s0 + rows) of every graph.  Pr points at the
// predecessor rows of those sources (row stride n, graph stride pstride).
//
// Level-synchronous tree accumulation, one workgroup (NT threads) per source
// s.  The predecessor row is a tree rooted at s; a vertex's gradient
// g[t] = G[s][t] + sum of its children's g is final once every deeper level
// is done, so:
//   1. depths by pointer jumping (ceil(log2 n) rounds: dep += dep[anc],
//      anc = anc[anc]);
//   2. vertices bucketed by depth (counting sort in LDS);
//   3. from the deepest level up: each vertex adds g[t] to its parent's g
//      (LDS float atomic); then the source's final g row goes to gout for
//      k_fcolumns, or, with gout null, each vertex adds g[t] to dW[P[s][t]][t]
//      with a global atomic as its level is done.
// Work O(n log n) per source with (log n + height) barriers, instead of a
// sort of all n keys followed by one lane walking n vertices.
// Dynamic LDS, 14 bytes per vertex + 8: par, dep, aux (int16: vertex ids
// and depths < 2^15), g (float), off (int32, n + 2).

template <int NT, int PER>
__global__ __launch_bounds__(NT) void k_fback(const float* __restrict__ D, const int32_t* __restrict__ Pr,
                                                        int64_t pstride, const float* __restrict__ G,
                                                        float* dW, int32_t n, int32_t s0, float* gout,
                                                        int64_t gstride, int32_t walk, int* __restrict__ err) {
  extern __shared__ __align__(16) unsigned char smem[];
  float* g = reinterpret_cast<float*>(smem);
  int32_t* off = reinterpret_cast<int32_t*>(g + n);
  int16_t* par = reinterpret_cast<int16_t*>(off + n + 2);
  int16_t* dep = par + n;
  int16_t* aux = dep + n;
  __shared__ int32_t s_h, s_changed[2];
  const int64_t gi = blockIdx.y;
  const int32_t lr = blockIdx.x;
  const int32_t src = s0 + lr;
  const int64_t nn = (int64_t)n * n;
  const float* Drow = D + gi * nn + (int64_t)src * n;
  const float* Grow = G + gi * nn + (int64_t)src * n;
  const int32_t* Prow = Pr + gi * pstride + (int64_t)lr * n;
  float* dWg = dW + gi * nn;
  const int tid = threadIdx.x;
  if (tid == 0) {
    s_h = 0;
    s_changed[0] = 0;
    s_changed[1] = 0;
  }
  for (int t = tid; t < n; t += NT) {
    const int32_t p = Prow[t];
    const bool reach = t != src && p >= 0 && Drow[t] < kFInf;
    par[t] = (int16_t)(reach ? p : -1);
    dep[t] = (int16_t)(t == src ? 0 : (reach ? 1 : -1));
    aux[t] = (int16_t)(reach ? p : -1);
    g[t] = Grow[t];
  }
  for (int d = tid; d < n + 2; d += NT) off[d] = 0;
  __syncthreads();
  // Shallow trees: when every reachable vertex is within kWalk parent links
  // of the source (checked by a bounded walk), each vertex adds its own G to
  // every ancestor below the source with LDS float atomics -- the same sums,
  // in another order, with two barriers instead of the pointer-jumping
  // rounds, the sort and one barrier per level.  Otherwise the level pass.
  constexpr int kWalk = 32;
  bool deep = !walk;
  for (int t = tid; t < n && !deep; t += NT) {
    int v = par[t];
    for (int h = 1; v >= 0 && v != src && h < kWalk; ++h) v = par[v];  // h: links followed
    deep = v >= 0 && v != src;  // par[t] < 0: the source or unreachable
    if (par[t] >= 0 && v < 0) deep = true;  // a broken chain: leave it to the level pass
  }
  if (!__syncthreads_or(deep)) {
    for (int t = tid; t < n; t += NT) {
      int v = par[t];
      if (v < 0) continue;
      const float x = Grow[t];
      while (v != src) {  // at most kWalk links, by the check above
        atomicAdd(&g[v], x);
        v = par[v];
      }
    }
    __syncthreads();
    if (!gout)
      for (int t = tid; t < n; t += NT)
        if (par[t] >= 0) atomicAdd(&dWg[(int64_t)par[t] * n + t], g[t]);
  } else {
    // 1. pointer jumping: after it, aux[t] == src and dep[t] = hops for every
    // reachable t.  Rounds are synchronous (all reads, barrier, all writes,
    // barrier); the "anything moved" flag of round r lives in slot r & 1 and
    // is cleared for round r + 1 only after every thread has passed round r's
    // first barrier, i.e. finished reading the slot of round r - 1.
    constexpr int kPer = PER;  // the host launches with n <= NT * PER
    for (int round = 0; round < 16; ++round) {
      int16_t nd[kPer], na[kPer];
      int changed = 0;
  #pragma unroll
      for (int e = 0; e < kPer; ++e) {
        const int t = tid + e * NT;
        nd[e] = 0;
        na[e] = -1;
        if (t < n) {
          const int a = aux[t];
          if (a >= 0 && a != src) {
            nd[e] = dep[a];
            na[e] = aux[a];
            changed = 1;
          }
        }
      }
      __syncthreads();
      if (changed) s_changed[round & 1] = 1;
  #pragma unroll
      for (int e = 0; e < kPer; ++e) {
        const int t = tid + e * NT;
        if (t < n && na[e] >= 0) {
          dep[t] = (int16_t)(dep[t] + nd[e]);
          aux[t] = na[e];
        }
      }
      __syncthreads();
      const int any = s_changed[round & 1];
      if (tid == 0) s_changed[(round + 1) & 1] = 0;
      if (!any) break;
    }
    // 16 doubling rounds bring every vertex of a tree of height < 2^16 to
    // the source.  A reachable vertex left elsewhere sits on a cycle or on a
    // chain that stops short of the source: the row is not a tree (zero or
    // fp32-absorbed weights can make recomputed predecessors point at each
    // other), and the gradient would be wrong, so the call reports it.
    bool broken = false;
    for (int t = tid; t < n; t += NT) broken |= par[t] >= 0 && aux[t] != src;
    if (__syncthreads_or(broken) && tid == 0) atomicOr(err, 1);
    // 2. counting sort by depth: off[d] counts level d, then exclusive prefix
    // sums (off[0] = 0 stays), then scatter into aux (order by level)
    for (int t = tid; t < n; t += NT) {
      const int d = dep[t];
      if (d > 0 && d <= n) {
        atomicAdd(&off[d], 1);
        atomicMax(&s_h, d);
      }
    }
    __syncthreads();
    const int H = s_h;
    if (tid == 0) {
      int run = 0;
      for (int d = 1; d <= H; ++d) {  // start of level d; serial: H is the tree height
        const int c = off[d];
        off[d] = run;
        run += c;
      }
    }
    __syncthreads();
    for (int t = tid; t < n; t += NT) {
      const int d = dep[t];
      if (d > 0 && d <= n) aux[atomicAdd(&off[d], 1)] = (int16_t)t;
    }
    __syncthreads();
    // now level d occupies [off[d - 1], off[d]) of aux (off[0] = 0)
    // 3. deepest level first
    for (int d = H; d >= 1; --d) {
      const int q0 = off[d - 1], q1 = off[d];
      for (int q = q0 + tid; q < q1; q += NT) {
        const int t = aux[q];
        const int p = par[t];
        const float gt = g[t];
        atomicAdd(&g[p], gt);
        if (!gout) atomicAdd(&dWg[(int64_t)p * n + t], gt);
      }
      __syncthreads();
    }
  }
  // column form: the final g of every vertex the levels above processed, 0
  // for the others (the source, unreachable vertices), one coalesced row per
  // source; k_fcolumns adds them into dW by predecessor.  gout may be dW
  // itself: no other block of this launch reads or writes row lr of it.
  if (gout) {
    float* row = gout + gi * gstride + (int64_t)lr * n;
    for (int t = tid; t < n; t += NT) {
      const int d = dep[t];
      row[t] = d > 0 && d <= n ? g[t] : 0.0f;
    }
  }
}

// dW[u][t] = sum of g_s[t] over the sources s of one chunk with P[s][t] = u
// (plus what earlier chunks left in dW unless `first`).  One workgroup per
// (CT target columns, graph) holds the whole column tile, n x CT floats, in
// LDS: it adds every source's g_s[t] of its columns there, at its
// predecessor, with an LDS float atomic, then writes each dW cell of the tile
// once.  Many sources reach a target through the same last edge, so the
// per-source global atomics on those cells were the backward's bottleneck;
// here the global traffic is one coalesced read of P and g and one write of
// dW.  Gb may alias dW (one chunk holding every source, first = 1): a
// workgroup reads only its own column tile of Gb, all of it before the
// barrier, and writes only that tile of dW after it.
template <int CT>
__global__ __launch_bounds__(1024) void k_fcolumns(const int32_t* __restrict__ P, int64_t pstride, const float* Gb,
                                                   int64_t gstride, int32_t rows, float* dW, int32_t n,
                                                   int32_t first) {
  constexpr int NT = 1024, U = 8;  // NT is a multiple of CT: a thread keeps one column
  extern __shared__ __align__(16) unsigned char smem[];
  float* acc = reinterpret_cast<float*>(smem);  // [u][c]
  const int c = threadIdx.x % CT, t = blockIdx.x * CT + c;
  const int64_t gi = blockIdx.y;
  const int64_t cells = (int64_t)n * CT;
  for (int64_t e = threadIdx.x; e < cells; e += NT) acc[e] = 0.0f;
  __syncthreads();
  const int32_t* Pg = P + gi * pstride;
  const float* Gg = Gb + gi * gstride;
  const int64_t items = (int64_t)rows * CT;
  // U loads in flight per thread before their LDS atomics
  for (int64_t e0 = threadIdx.x; e0 < items; e0 += (int64_t)NT * U) {
    int32_t p[U];
    float v[U];
#pragma unroll
    for (int k = 0; k < U; ++k) {
      const int64_t e = e0 + (int64_t)k * NT;
      const bool in = e < items && t < n;
      const int64_t at = (e / CT) * n + t;
      p[k] = in ? Pg[at] : -1;
      v[k] = in ? Gg[at] : 0.0f;
    }
#pragma unroll
    for (int k = 0; k < U; ++k)
      if (p[k] >= 0 && p[k] < n) atomicAdd(&acc[p[k] * CT + c], v[k]);
  }
  __syncthreads();
  if (t >= n) return;
  float* dWg = dW + gi * (int64_t)n * n;
  for (int64_t e = threadIdx.x; e < cells; e += NT) {
    float* o = dWg + (e / CT) * n + t;
    *o = first ? acc[e] : *o + acc[e];
  }
}

// Columns per workgroup: the largest power of two in [4, 64] with n x CT
// floats within 128 KiB of LDS.
static int fcolumns_ct(int32_t n) {
  int ct = 64;
  while (ct > 4 && (int64_t)n * ct * 4 > (128 << 10)) ct >>= 1;
  return ct;
}

template <int CT>
static hipError_t fcolumns_launch(const int32_t* P, int64_t pstride, const float* Gb, int64_t gstride, int32_t rows,
                                  float* dW, int32_t n, int32_t batch, bool first, hipStream_t st) {
  const size_t lds = (size_t)n * CT * sizeof(float);
  if (lds > 65536) {
    const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_fcolumns<CT>),
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
  }
  k_fcolumns<CT><<<dim3((n + CT - 1) / CT, batch), 1024, lds, st>>>(P, pstride, Gb, gstride, rows, dW, n,
                                                                     first ? 1 : 0);
  return hipGetLastError();
}

static hipError_t fcolumns(const int32_t* P, int64_t pstride, const float* Gb, int64_t gstride, int32_t rows, float* dW,
                           int32_t n, int32_t batch, bool first, hipStream_t st) {
  switch (fcolumns_ct(n)) {
    case 64: return fcolumns_launch<64>(P, pstride, Gb, gstride, rows, dW, n, batch, first, st);
    case 32: return fcolumns_launch<32>(P, pstride, Gb, gstride, rows, dW, n, batch, first, st);
    case 16: return fcolumns_launch<16>(P, pstride, Gb, gstride, rows, dW, n, batch, first, st);
    case 8: return fcolumns_launch<8>(P, pstride, Gb, gstride, rows, dW, n, batch, first, st);
    default: return fcolumns_launch<4>(P, pstride, Gb, gstride, rows, dW, n, batch, first, st);
  }
}

// FWX_DIFF_WALK=0: every source's tree is accumulated level by level, also
// when it is shallow (read on every call).
static bool diff_back_walk() {
  const char* e = std::getenv("FWX_DIFF_WALK");
  return !(e && e[0] == '0');
}

// FWX_DIFF_BACK=0: each source adds its contributions to dW with one global
// atomic per vertex (read on every call).
static bool diff_back_columns() {
  const char* e = std::getenv("FWX_DIFF_BACK");
  return !(e && e[0] == '0');
}

// Product kernels: FWX_DIFF_REST=0 selects the first versions (k_frest,
// k_fpanel: scalar LDS reads, single-buffered), bit-identical to the default
// k_frest4 / k_fpanel4 (read on every call).
static bool diff_rest_fast() {
  const char* e = std::getenv("FWX_DIFF_REST");
  return !(e && e[0] == '0');
}

// FWX_DIFF_DIAG=0 selects the first diagonal closure (k_fdiag: 256
// threads, one k step per barrier), bit-identical to k_fdiag2.
static bool diff_diag_fast() {
  const char* e = std::getenv("FWX_DIFF_DIAG");
  return !(e && e[0] == '0');
}

// FWX_LEAD_PRIO=0 keeps the side stream's kernels (band, closure, panels)
// at the default wave priority.
static int diff_side_prio() {
  const char* e = std::getenv("FWX_LEAD_PRIO");
  return (e && e[0] == '0') ? 0 : 1;
}

template <bool PRED>
void frest_launch(float* D, int32_t* P, int32_t n, int32_t c, int32_t N, int32_t mode, int32_t blocks,
                  int32_t batch, int32_t prio, hipStream_t st) {
  const bool vec = (n % 4 == 0) && ((reinterpret_cast<uintptr_t>(D) & 15u) == 0) &&
                   (!PRED || (reinterpret_cast<uintptr_t>(P) & 15u) == 0);
  if (!diff_rest_fast())
    k_frest<PRED><<<dim3(blocks, batch), 256, 0, st>>>(D, P, n, c, N, mode);
  else if (vec)
    k_frest4<PRED, true><<<dim3(blocks, batch), 256, 0, st>>>(D, P, n, c, N, mode, prio);
  else
    k_frest4<PRED, false><<<dim3(blocks, batch), 256, 0, st>>>(D, P, n, c, N, mode, prio);
}

template <bool PRED>
int lead_round(float* D, int32_t* P, int32_t n, int32_t c, int32_t N, int32_t batch, int32_t prio,
               hipStream_t st) {
  if (diff_diag_fast()) {
    k_fdiag2<PRED><<<dim3(1, batch), 1024, 0, st>>>(D, P, n, c, prio);
    FWX_LAUNCHED("k_fdiag2");
  } else {
    k_fdiag<PRED><<<dim3(1, batch), 256, 0, st>>>(D, P, n, c);
    FWX_LAUNCHED("k_fdiag");
  }
  if (N > 1) {
    const bool vec = (n % 4 == 0) && ((reinterpret_cast<uintptr_t>(D) & 15u) == 0) &&
                     (!PRED || (reinterpret_cast<uintptr_t>(P) & 15u) == 0);
    if (!diff_rest_fast())
      k_fpanel<PRED><<<dim3(2 * (N - 1), batch), 256, 0, st>>>(D, P, n, c, N, prio);
    else if (vec)
      k_fpanel4<PRED, true><<<dim3(2 * (N - 1), batch), 256, 0, st>>>(D, P, n, c, N, prio);
    else
      k_fpanel4<PRED, false><<<dim3(2 * (N - 1), batch), 256, 0, st>>>(D, P, n, c, N, prio);
    FWX_LAUNCHED("k_fpanel");
  }
  return FWX_OK;
}

// Three-phase blocked FW over 64-tiles.  Schedules (FWX_LOOKAHEAD, read on
// every call; all give the same bits, since every tile sees the stages in
// order):
//  0  one stream: lead(c), rest(c), lead(c+1), ...
//  1  look-ahead: stage c's rest pass first updates the band that stage
//     c+1's closure and panels read (row and column c+1), then the side
//     stream runs lead(c+1) while the main stream updates the other tiles;
//     the main stream waits for it before stage c+1's band.
//  2  side-band (default): the side stream runs band(c) and then lead(c+1);
//     the main stream runs only the other rest tiles of each stage.  band(c)
//     and lead(c+1) write only row/column c+1 tiles; rest_other(c) reads
//     row/column c tiles and writes tiles outside rows/columns c and c+1.
//     rest_other(c) waits for lead(c) (event L); band(c+1) waits for
//     rest_other(c) (event R), which wrote the row/column c+2 tiles it
//     updates next.
static int diff_schedule() {
  const char* e = std::getenv("FWX_LOOKAHEAD");
  return (e && e[0] >= '0' && e[0] <= '2') ? e[0] - '0' : 2;
}

template <bool PRED>
int forward_sideband(float* D, int32_t* P, int32_t n, int32_t N, int32_t batch, hipStream_t st, hipStream_t side,
                     hipEvent_t evR, hipEvent_t evL) {
  FWX_HIP(hipEventRecord(evR, st));  // the side stream starts after the caller's prior work
  FWX_HIP(hipStreamWaitEvent(side, evR, 0));
  const int32_t prio = diff_side_prio();
  if (int rc = lead_round<PRED>(D, P, n, 0, N, batch, prio, side)) return rc;
  FWX_HIP(hipEventRecord(evL, side));
  for (int c = 0; c < N; ++c) {
    FWX_HIP(hipStreamWaitEvent(st, evL, 0));
    if (c + 1 < N) {
      frest_launch<PRED>(D, P, n, c, N, 1, 2 * N - 3, batch, prio, side);
      FWX_LAUNCHED("k_frest");
      if (int rc = lead_round<PRED>(D, P, n, c + 1, N, batch, prio, side)) return rc;
      FWX_HIP(hipEventRecord(evL, side));
      frest_launch<PRED>(D, P, n, c, N, 2, (N - 2) * (N - 2), batch, 0, st);
      FWX_LAUNCHED("k_frest");
      FWX_HIP(hipEventRecord(evR, st));
      FWX_HIP(hipStreamWaitEvent(side, evR, 0));
    } else {
      frest_launch<PRED>(D, P, n, c, N, 0, (N - 1) * (N - 1), batch, 0, st);
      FWX_LAUNCHED("k_frest");
    }
  }
  return FWX_OK;
}

// Stage width in k-tiles for N 64-wide tiles: 4 from N = 128 (n = 8192),
// 2 from N = 48, else 1 -- forward ms measured (best of 3 x 4 calls):
// n=8192 64.1 / 57.7 / 52.9 (kappa 1 / 2 / 4); n=4096 8.05 / 7.44 / 8.15;
// n=3072 6.80 / 6.54; n=2048 4.16 / 4.23; n=512 (batch 64) 1.24 / 1.39.
// FWX_DIFF_KAPPA=1..8 overrides (read on every call).
static int diff_kappa(int N) {
  if (const char* e = std::getenv("FWX_DIFF_KAPPA")) {
    const int k = std::atoi(e);
    if (k >= 1 && k <= 8) return k;
  }
  return N >= 128 ? 4 : N >= 48 ? 2 : 1;
}

template <bool PRED>
static bool fvec(const float* D, const int32_t* P, int32_t n) {
  return (n % 4 == 0) && ((reinterpret_cast<uintptr_t>(D) & 15u) == 0) &&
         (!PRED || (reinterpret_cast<uintptr_t>(P) & 15u) == 0);
}

// The lead rounds of stage [kb, kb+Bst): closure, panels and cross of each
// of its k-tiles.
template <bool PRED>
int lead_stage_ml(float* D, int32_t* P, int32_t n, int32_t kb, int32_t Bst, int32_t N, int32_t batch,
                  int32_t prio, hipStream_t st) {
  const bool vec = fvec<PRED>(D, P, n);
  for (int l = 0; l < Bst; ++l) {
    const int c = kb + l;
    if (int rc = lead_round<PRED>(D, P, n, c, N, batch, prio, st)) return rc;
    if (Bst > 1) {
      const int cnt = (Bst - 1) * (2 * N - Bst - 1);
      if (vec)
        k_fcross<PRED, true><<<dim3(cnt, batch), 256, 0, st>>>(D, P, n, c, kb, Bst, N, prio);
      else
        k_fcross<PRED, false><<<dim3(cnt, batch), 256, 0, st>>>(D, P, n, c, kb, Bst, N, prio);
      FWX_LAUNCHED("k_fcross");
    }
  }
  return FWX_OK;
}

template <bool PRED>
void frest_ml_launch(float* D, int32_t* P, int32_t n, int32_t e0, int32_t en, int32_t fn, int32_t N, int32_t mode,
                     int32_t batch, int32_t prio, hipStream_t st) {
  const int m = N - en;
  const int blocks = mode == 0 ? m * m : mode == 1 ? fn * m + (m - fn) * fn : (m - fn) * (m - fn);
  if (blocks <= 0) return;
  if (fvec<PRED>(D, P, n))
    k_frest_ml<PRED, true><<<dim3(blocks, batch), 256, 0, st>>>(D, P, n, e0, en, fn, N, mode, prio);
  else
    k_frest_ml<PRED, false><<<dim3(blocks, batch), 256, 0, st>>>(D, P, n, e0, en, fn, N, mode, prio);
}

// Multi-layer forward: one stream (mode 0) or side-band (otherwise; the
// same read/write-set argument as the one-k-tile schedule, with stages in
// place of k-tiles).
template <bool PRED>
int forward_ml(float* D, int32_t* P, int32_t n, int32_t N, int32_t kappa, int32_t batch, int mode,
               hipStream_t st) {
  if (mode == 0) {
    for (int kb = 0; kb < N; kb += kappa) {
      const int Bst = std::min(kappa, N - kb);
      if (int rc = lead_stage_ml<PRED>(D, P, n, kb, Bst, N, batch, 0, st)) return rc;
      frest_ml_launch<PRED>(D, P, n, kb, Bst, 0, N, 0, batch, 0, st);
      FWX_LAUNCHED("k_frest_ml");
    }
    return FWX_OK;
  }
  hipStream_t side;
  if (int rc = side_stream(&side)) return rc;
  hipEvent_t evR, evL;
  FWX_HIP(hipEventCreateWithFlags(&evR, event_flags(false)));
  hipError_t e = hipEventCreateWithFlags(&evL, event_flags(false));
  if (e != hipSuccess) {
    (void)hipEventDestroy(evR);
    return hip_fail(e, "hipEventCreateWithFlags", __FILE__, __LINE__);
  }
  const int32_t prio = diff_side_prio();
  auto body = [&]() -> int {
    FWX_HIP(hipEventRecord(evR, st));
    FWX_HIP(hipStreamWaitEvent(side, evR, 0));
    if (int rc = lead_stage_ml<PRED>(D, P, n, 0, std::min(kappa, N), N, batch, prio, side)) return rc;
    FWX_HIP(hipEventRecord(evL, side));
    for (int kb = 0; kb < N; kb += kappa) {
      const int Bst = std::min(kappa, N - kb);
      const int kb2 = kb + kappa;
      FWX_HIP(hipStreamWaitEvent(st, evL, 0));
      if (kb2 < N) {
        const int B2 = std::min(kappa, N - kb2);
        frest_ml_launch<PRED>(D, P, n, kb, Bst, B2, N, 1, batch, prio, side);
        FWX_LAUNCHED("k_frest_ml");
        if (int rc = lead_stage_ml<PRED>(D, P, n, kb2, B2, N, batch, prio, side)) return rc;
        FWX_HIP(hipEventRecord(evL, side));
        frest_ml_launch<PRED>(D, P, n, kb, Bst, B2, N, 2, batch, 0, st);
        FWX_LAUNCHED("k_frest_ml");
        FWX_HIP(hipEventRecord(evR, st));
        FWX_HIP(hipStreamWaitEvent(side, evR, 0));
      } else {
        frest_ml_launch<PRED>(D, P, n, kb, Bst, 0, N, 0, batch, 0, st);
        FWX_LAUNCHED("k_frest_ml");
      }
    }
    return FWX_OK;
  };
  const int rc = body();
  (void)hipEventDestroy(evR);
  (void)hipEventDestroy(evL);
  return rc;
}

template <bool PRED>
int forward_impl(const float* W, float* D, int32_t* P, int32_t n, int32_t batch, hipStream_t st) {
  const int64_t total = (int64_t)n * n * batch;
  const int fillblocks = (int)std::min<int64_t>((total + 255) / 256, 65536);
  k_finit<PRED><<<fillblocks, 256, 0, st>>>(W, D, P, n, total);
  FWX_LAUNCHED("k_finit");
  const int N = (n + FT - 1) / FT;
  const int mode = diff_schedule();
  const int kappa = diff_kappa(N);
  if (kappa > 1 && N > kappa) return forward_ml<PRED>(D, P, n, N, kappa, batch, mode, st);
  if (N < 3 || mode == 0) {
    for (int c = 0; c < N; ++c) {
      if (int rc = lead_round<PRED>(D, P, n, c, N, batch, 0, st)) return rc;
      if (N > 1) {
        frest_launch<PRED>(D, P, n, c, N, 0, (N - 1) * (N - 1), batch, 0, st);
        FWX_LAUNCHED("k_frest");
      }
    }
    return FWX_OK;
  }
  hipStream_t side;
  if (int rc = side_stream(&side)) return rc;
  hipEvent_t evA, evB;
  FWX_HIP(hipEventCreateWithFlags(&evA, event_flags(false)));
  hipError_t e = hipEventCreateWithFlags(&evB, event_flags(false));
  if (e != hipSuccess) {
    (void)hipEventDestroy(evA);
    return hip_fail(e, "hipEventCreateWithFlags", __FILE__, __LINE__);
  }
  int rc = FWX_OK;
  if (mode == 2) {
    rc = forward_sideband<PRED>(D, P, n, N, batch, st, side, evA, evB);
    (void)hipEventDestroy(evA);
    (void)hipEventDestroy(evB);
    return rc;
  }
  rc = lead_round<PRED>(D, P, n, 0, N, batch, 0, st);
  for (int c = 0; c < N && rc == FWX_OK; ++c) {
    if (c + 1 < N) {
      frest_launch<PRED>(D, P, n, c, N, 1, 2 * N - 3, batch, 0, st);
      e = hipGetLastError();
      if (e == hipSuccess) e = hipEventRecord(evA, st);
      if (e == hipSuccess) e = hipStreamWaitEvent(side, evA, 0);
      if (e != hipSuccess) {
        rc = hip_fail(e, "diff look-ahead band", __FILE__, __LINE__);
        break;
      }
      if ((rc = lead_round<PRED>(D, P, n, c + 1, N, batch, diff_side_prio(), side))) break;
      e = hipEventRecord(evB, side);
      if (e == hipSuccess) {
        frest_launch<PRED>(D, P, n, c, N, 2, (N - 2) * (N - 2), batch, 0, st);
        e = hipGetLastError();
      }
      if (e == hipSuccess) e = hipStreamWaitEvent(st, evB, 0);
      if (e != hipSuccess) rc = hip_fail(e, "diff look-ahead rest", __FILE__, __LINE__);
    } else {
      frest_launch<PRED>(D, P, n, c, N, 0, (N - 1) * (N - 1), batch, 0, st);
      e = hipGetLastError();
      if (e != hipSuccess) rc = hip_fail(e, "k_frest", __FILE__, __LINE__);
    }
  }
  (void)hipEventDestroy(evA);
  (void)hipEventDestroy(evB);
  return rc;
}

constexpr int32_t kDiffMaxN = 8192;  // backward keeps 14 bytes per vertex in LDS

int check_diff(const void* a, const void* b, int32_t n, int32_t batch) {
  if (!a || !b) return fail(FWX_ERR_INVALID_ARG, "null matrix pointer");
  if (n <= 0 || n > kDiffMaxN) return fail(FWX_ERR_INVALID_ARG, "differentiable FW needs 1 <= n <= 8192");
  if (batch <= 0 || batch > 65535) return fail(FWX_ERR_INVALID_ARG, "batch must be in [1, 65535]");
  return FWX_OK;
}

}  // namespace

}  // namespace fwx

using namespace fwx;

// Workspace of fwx_diff_backward_noalloc: 16 MiB per device, in the code
// object's own device memory (allocated when the library is loaded on a
// device, held until the process ends).  Its name in host code is a
// host-side stand-in, not a device address: host code uses only the
// per-device address from scratch_address().  Every call on one device
// shares it, so calls that overlap on one device race.
constexpr size_t kScratchBytes = 16u << 20;
static __device__ __align__(256) int32_t scratch[kScratchBytes / sizeof(int32_t)];

static std::mutex g_scratch_mu;
static int32_t* g_scratch_ptr[64] = {};

// The current device's address of `scratch`, looked up once per device
// (with that device current) on the first call there.
static int scratch_address(int32_t** out) {
  int dev = 0;
  FWX_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) return fail(FWX_ERR_UNSUPPORTED, "device index above 63");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    int32_t* ptr = nullptr;
    FWX_HIP(hipGetSymbolAddress(reinterpret_cast<void**>(&ptr), HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = ptr;
  }
  *out = g_scratch_ptr[dev];
  return FWX_OK;
}

// The backward of fwx_diff_backward.  Recomputed predecessors (d_P null)
// need a workspace of rows x n x batch int32 (and as many floats of g rows in
// the column form), taken from hipMallocAsync with rows = 256 when
// `workspace` is null, else from `workspace` (workspace_bytes long) with as many
// sources per chunk as fit, at most 256.  Every workspace cell a kernel
// reads is written earlier in the same chunk, so it needs no clearing.
// Ends with `st` synchronised; a predecessor row that is not a tree rooted
// at its source returns FWX_ERR_INVALID_ARG.
static int backward_impl(const float* d_W, const float* d_D, const int32_t* d_P, const float* d_G, float* d_dW,
                         int32_t n, int32_t batch, hipStream_t st, int32_t* workspace, size_t workspace_bytes) {
  if (int rc = check_diff(d_D, d_G, n, batch)) return rc;
  if (!d_dW) return fail(FWX_ERR_INVALID_ARG, "null gradient output");
  if (!d_P && !d_W) return fail(FWX_ERR_INVALID_ARG, "recomputing predecessors needs W");
  const size_t lds = (size_t)n * 14 + 8;
  // Column form (default): k_fback writes each source's g row, k_fcolumns
  // adds the rows into dW by predecessor.  With stored predecessors the g
  // rows go into dW itself (one chunk of every source); recomputed
  // predecessors come in chunks of at most 256 sources with a g buffer
  // beside them.  FWX_DIFF_BACK=0 zeroes dW and adds to it with global
  // atomics.
  const bool columns = diff_back_columns();
  const int32_t walk = diff_back_walk() ? 1 : 0;
  int32_t rows = std::min<int32_t>(256, n);
  const size_t per_row = (size_t)n * batch * (columns ? 8 : 4);
  if (!d_P && workspace) {
    const size_t fit = workspace_bytes / per_row;
    if (fit == 0)
      return fail(FWX_ERR_INVALID_ARG, "n x batch too large for the workspace: one source row of every graph "
                                       "needs " + std::to_string(per_row) + " bytes, the workspace holds " +
                                       std::to_string(workspace_bytes));
    rows = (int32_t)std::min<size_t>((size_t)rows, fit);
  }
  DeviceFlag fl;  // set by k_fback when a predecessor row is not a tree
  if (int rc = flag_acquire(st, &fl)) return rc;
  int rc = FWX_OK;
  if (!columns) {
    const int64_t total = (int64_t)n * n * batch;
    k_ffill<<<(int)std::min<int64_t>((total + 255) / 256, 65536), 256, 0, st>>>(d_dW, total);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) rc = hip_fail(e, "k_ffill", __FILE__, __LINE__);
  }
  // one workgroup per source: 256 threads up to n = 1024, 512 up to 4096,
  // 1024 above (each thread holds at most PER vertices in the jumping rounds)
  auto back = [&](const int32_t* pr, int64_t pstride, int32_t nrows, int32_t s0, float* gb,
                  int64_t gstride) -> hipError_t {
    const dim3 grid(nrows, batch);
    if (n <= 1024) {
      k_fback<256, 4><<<grid, 256, lds, st>>>(d_D, pr, pstride, d_G, d_dW, n, s0, gb, gstride, walk, fl.f);
    } else if (n <= 4096) {
      k_fback<512, 8><<<grid, 512, lds, st>>>(d_D, pr, pstride, d_G, d_dW, n, s0, gb, gstride, walk, fl.f);
    } else {
      if (lds > 65536) {
        const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_fback<1024, 8>),
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
      }
      k_fback<1024, 8><<<grid, 1024, lds, st>>>(d_D, pr, pstride, d_G, d_dW, n, s0, gb, gstride, walk, fl.f);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || !gb) return e;
    return fcolumns(pr, pstride, gb, gstride, nrows, d_dW, n, batch, s0 == 0, st);
  };
  if (rc) {
  } else if (d_P) {
    const int64_t nn = (int64_t)n * n;
    const hipError_t e = back(d_P, nn, n, 0, columns ? d_dW : nullptr, nn);
    if (e != hipSuccess) rc = hip_fail(e, "k_fback", __FILE__, __LINE__);
  } else {
    const size_t cnt = (size_t)rows * n * batch;
    int32_t* ws = workspace;
    if (!ws) {
      const hipError_t ea = hipMallocAsync(reinterpret_cast<void**>(&ws), cnt * (columns ? 8 : 4), st);
      if (ea != hipSuccess) {
        ws = nullptr;
        rc = hip_fail(ea, "hipMallocAsync", __FILE__, __LINE__);
      }
    }
    float* gb = columns && ws ? reinterpret_cast<float*>(ws + cnt) : nullptr;
    for (int32_t s0 = 0; s0 < n && rc == FWX_OK; s0 += rows) {
      const int32_t r = std::min(rows, n - s0);
      const int nt = (n + FT - 1) / FT;
      k_fpred<<<dim3(((r + FT - 1) / FT) * nt, batch), 256, 0, st>>>(d_W, d_D, ws, n, s0, r);
      hipError_t e = hipGetLastError();
      if (e == hipSuccess) e = back(ws, (int64_t)r * n, r, s0, gb, (int64_t)r * n);
      if (e != hipSuccess) rc = hip_fail(e, "k_fpred/k_fback", __FILE__, __LINE__);
    }
    if (ws && !workspace) (void)hipFreeAsync(ws, st);
  }
  int bad = 0;
  const int rc2 = flag_release(st, &fl, rc ? nullptr : &bad);
  if (rc) return rc;
  if (rc2) return rc2;
  if (bad)
    return fail(FWX_ERR_INVALID_ARG,
                "a predecessor row is not a tree rooted at its source (a cycle or a broken chain: zero weights, "
                "or weights too small to change an fp32 distance); dW is undefined");
  return FWX_OK;
}

extern "C" {

int fwx_diff_forward(const float* d_W, float* d_D, int32_t* d_P, int32_t n, int32_t batch, void* stream) {
  clear_error();
  if (int rc = check_diff(d_W, d_D, n, batch)) return rc;
  hipStream_t st = as_stream(stream);
  const int rc = d_P ? forward_impl<true>(d_W, d_D, d_P, n, batch, st)
                     : forward_impl<false>(d_W, d_D, nullptr, n, batch, st);
  if (rc) return rc;
  FWX_HIP(hipStreamSynchronize(st));
  return FWX_OK;
}

int fwx_diff_backward(const float* d_W, const float* d_D, const int32_t* d_P, const float* d_G, float* d_dW,
                      int32_t n, int32_t batch, void* stream) {
  clear_error();
  return backward_impl(d_W, d_D, d_P, d_G, d_dW, n, batch, as_stream(stream), nullptr, 0);
}

int fwx_diff_backward_noalloc(const float* d_W, const float* d_D, const float* d_G, float* d_dW, int32_t n,
                              int32_t batch, void* stream) {
  clear_error();
  if (int rc = check_diff(d_D, d_G, n, batch)) return rc;
  if (!d_W) return fail(FWX_ERR_INVALID_ARG, "recomputing predecessors needs W");
  const size_t per_row = (size_t)n * batch * (diff_back_columns() ? 8 : 4);
  if (per_row > kScratchBytes)
    return fail(FWX_ERR_INVALID_ARG, "fwx_diff_backward_noalloc: n x batch = " + std::to_string((size_t)n * batch) +
                                         " exceeds the 16 MiB workspace (at most 2097152 in the column form)");
  int32_t* ws = nullptr;
  if (int rc = scratch_address(&ws)) return rc;
  return backward_impl(d_W, d_D, nullptr, d_G, d_dW, n, batch, as_stream(stream), ws, kScratchBytes);
}

}  // extern "C"
