// fwx_tile.hpp -- register-blocked min-plus tile product for gfx950.
//
//   C[BM x BN] <- min(C, A[BM x K] (x) B[K x BN])      ((x) = min-plus product)
//
// One 256-thread workgroup (four wave64s, one per SIMD) owns the C block in
// VGPRs: thread (tx, ty) = (tid % 16, tid / 16) holds cells
// (ty + 16 r, tx + 16 c) for r < BM/16, c < BN/16.  The interleaved mapping
// makes the LDS reads conflict-free: a wave's 16 tx lanes read 16
// consecutive B pairs (128 B) and its 4 ty rows hit 4 distinct 16-byte bank
// groups of A (row stride KC+4 dwords), the rest are broadcasts.
//
// K is walked in chunks of KC, double-buffered through LDS with one barrier
// per chunk.  A is staged row-major (k contiguous), B pair-interleaved
// (B[k][j], B[k+1][j] adjacent), so both operands of a k pair are one 64-bit
// register each -- the layout the pk64 form needs (see fwx_device.hpp).
// While staging, each wave ballots whether any operand of the chunk lies
// outside [0, 2^30); the flag rides on the chunk barrier and selects the
// pk64 or the min3 form for the whole chunk, uniformly across the block.
//
// Out-of-range operand cells (row/col/k >= n) are staged as INF and C cells
// outside the matrix are never stored, so ragged n needs no padding buffer.
#pragma once

#include "fwx_device.hpp"

// Unroll of the k-quad loop inside a chunk (register pressure vs. LDS
// latency hiding).
#ifndef FWX_QUNROLL
#define FWX_QUNROLL 1
#endif

namespace fwx {

// A rectangular window of a row-major int32 matrix.  `p` points at the
// window's (0, 0) cell (uniform across the block, so it lives in SGPRs);
// element (r, c) is p[r * ld + c] -- a 32-bit offset, since a window spans
// at most a few thousand rows -- and is in range iff r < rlim && c < clim.
struct View {
  int32_t* p;
  int32_t ld;
  int32_t rlim, clim;
  __device__ __forceinline__ bool in(int32_t r, int32_t c) const { return r < rlim && c < clim; }
  __device__ __forceinline__ int32_t* at(int32_t r, int32_t c) const { return p + (r * ld + c); }
};

__device__ __forceinline__ View make_view(int32_t* D, int32_t n, int32_t r0, int32_t c0) {
  return View{D + ((int64_t)r0 * n + c0), n, n - r0, n - c0};
}

// VEC: 16-byte global loads (needs ld % 4 == 0, 16-byte aligned base and
// c0 % 4 == 0, so a 4-wide group is wholly in or wholly out of range).
template <int BM, int BN, int KC, bool VEC>
struct TileProduct {
  static constexpr int NT = 256;
  static constexpr int TM = BM / 16;
  static constexpr int TN = BN / 16;
  static constexpr int LDA = KC + 4;
  static_assert(BM % 16 == 0 && BN % 16 == 0, "block must be a multiple of 16");
  static_assert(KC % 4 == 0, "chunk must hold whole k quads");
  // A staging: KC/4 int4 groups per row.
  static constexpr int AG = KC / 4;
  static constexpr int AROWS = NT / AG;  // rows per pass
  static constexpr int APASS = (BM + AROWS - 1) / AROWS;
  // B staging: units of (k pair, 4 columns).
  static constexpr int BU = (KC / 2) * (BN / 4);
  static constexpr int BPASS = (BU + NT - 1) / NT;

  struct Smem {
    int32_t a[2][BM * LDA];
    int32_t b[2][(KC / 2) * BN * 2];
    int flag[2][4];
  };

  int4 ra[APASS];
  int4 rb0[BPASS], rb1[BPASS];
  bool bad;

  __device__ __forceinline__ static int4 load4(const View& v, int32_t r, int32_t c) {
    if (VEC) {
      if (v.in(r, c)) return *reinterpret_cast<const int4*>(v.at(r, c));
      return make_int4(kInf, kInf, kInf, kInf);
    } else {
      int4 x;
      x.x = v.in(r, c + 0) ? *v.at(r, c + 0) : kInf;
      x.y = v.in(r, c + 1) ? *v.at(r, c + 1) : kInf;
      x.z = v.in(r, c + 2) ? *v.at(r, c + 2) : kInf;
      x.w = v.in(r, c + 3) ? *v.at(r, c + 3) : kInf;
      return x;
    }
  }

  __device__ __forceinline__ static uint32_t orbits(int4 x) {
    return (uint32_t)x.x | (uint32_t)x.y | (uint32_t)x.z | (uint32_t)x.w;
  }

  // global -> registers for the chunk starting at k offset kc0.  When the
  // whole chunk is inside the matrix (the common case) every address is a
  // uniform row base (SGPRs) plus one per-thread offset (one VGPR) plus an
  // immediate; the bounds-checked path handles the ragged edge.
  __device__ __forceinline__ void fetch(const View& A, const View& B, int32_t kc0) {
    const int tid = threadIdx.x;
    uint32_t o = 0;
    static_assert(NT % (BN / 4) == 0, "B staging rows must be uniform per pass");
    constexpr int BROWS = NT / (BN / 4);  // k pairs per B pass
    const bool full = VEC && A.rlim >= BM && A.clim >= kc0 + KC && B.rlim >= kc0 + KC && B.clim >= BN;
    if (full) {
      const int32_t* pa = A.p + kc0;
      const int32_t oa = (tid / AG) * A.ld + (tid % AG) * 4;
#pragma unroll
      for (int p = 0; p < APASS; ++p) {
        if (BM % AROWS == 0 || p * AROWS + tid / AG < BM) {
          ra[p] = *reinterpret_cast<const int4*>(pa + (p * AROWS) * A.ld + oa);
          o |= orbits(ra[p]);
        }
      }
      const int32_t* pb = B.p + kc0 * B.ld;
      const int32_t ob = (tid / (BN / 4)) * 2 * B.ld + (tid % (BN / 4)) * 4;
#pragma unroll
      for (int u = 0; u < BPASS; ++u) {
        if (BU % NT == 0 || u * NT + tid < BU) {
          const int32_t* q = pb + (u * BROWS * 2) * B.ld + ob;
          rb0[u] = *reinterpret_cast<const int4*>(q);
          rb1[u] = *reinterpret_cast<const int4*>(q + B.ld);
          o |= orbits(rb0[u]) | orbits(rb1[u]);
        }
      }
    } else {
#pragma unroll
      for (int p = 0; p < APASS; ++p) {
        const int r = p * AROWS + tid / AG;
        const int k = (tid % AG) * 4;
        if (BM % AROWS == 0 || r < BM) {
          ra[p] = load4(A, r, kc0 + k);
          o |= orbits(ra[p]);
        }
      }
#pragma unroll
      for (int u = 0; u < BPASS; ++u) {
        const int idx = u * NT + tid;
        if (BU % NT == 0 || idx < BU) {
          const int pr = idx / (BN / 4);
          const int j = (idx % (BN / 4)) * 4;
          rb0[u] = load4(B, kc0 + 2 * pr, j);
          rb1[u] = load4(B, kc0 + 2 * pr + 1, j);
          o |= orbits(rb0[u]) | orbits(rb1[u]);
        }
      }
    }
    bad = (o & kPk64Mask) != 0u;
  }

  // registers -> LDS buffer `buf`
  __device__ __forceinline__ void stash(Smem& s, int buf) {
    const int tid = threadIdx.x;
#pragma unroll
    for (int p = 0; p < APASS; ++p) {
      const int r = p * AROWS + tid / AG;
      const int k = (tid % AG) * 4;
      if (BM % AROWS == 0 || r < BM)
        *reinterpret_cast<int4*>(&s.a[buf][r * LDA + k]) = ra[p];
    }
#pragma unroll
    for (int u = 0; u < BPASS; ++u) {
      const int idx = u * NT + tid;
      if (BU % NT == 0 || idx < BU) {
        const int pr = idx / (BN / 4);
        const int j = (idx % (BN / 4)) * 4;
        int32_t* dst = &s.b[buf][(pr * BN + j) * 2];
        *reinterpret_cast<int4*>(dst) = make_int4(rb0[u].x, rb1[u].x, rb0[u].y, rb1[u].y);
        *reinterpret_cast<int4*>(dst + 4) = make_int4(rb0[u].z, rb1[u].z, rb0[u].w, rb1[u].w);
      }
    }
    publish_wave_flag<4>(s.flag[buf], bad);
  }

  template <bool PK64>
  __device__ __forceinline__ static void compute(const Smem& s, int buf, int32_t (&acc)[TM][TN]) {
    const int tx = threadIdx.x & 15;
    const int ty = threadIdx.x >> 4;
    const int32_t* As = s.a[buf];
    const uint64_t* Bs = reinterpret_cast<const uint64_t*>(s.b[buf]);
#pragma unroll FWX_QUNROLL
    for (int q = 0; q < KC / 4; ++q) {
      int4 a[TM];
      uint64_t b0[TN], b1[TN];
#pragma unroll
      for (int r = 0; r < TM; ++r)
        a[r] = *reinterpret_cast<const int4*>(&As[(ty + 16 * r) * LDA + 4 * q]);
#pragma unroll
      for (int c = 0; c < TN; ++c) {
        b0[c] = Bs[(2 * q) * BN + tx + 16 * c];
        b1[c] = Bs[(2 * q + 1) * BN + tx + 16 * c];
      }
#pragma unroll
      for (int r = 0; r < TM; ++r) {
        const uint64_t a0 = pack2(a[r].x, a[r].y);
        const uint64_t a1 = pack2(a[r].z, a[r].w);
#pragma unroll
        for (int c = 0; c < TN; ++c) {
          if (PK64) {
            acc[r][c] = relax2_pk64(acc[r][c], a0, b0[c]);
            acc[r][c] = relax2_pk64(acc[r][c], a1, b1[c]);
          } else {
            acc[r][c] = relax2_min3(acc[r][c], a0, b0[c]);
            acc[r][c] = relax2_min3(acc[r][c], a1, b1[c]);
          }
        }
      }
    }
  }

  // acc <- min(acc, A (x) B) over k in [0, depth), depth a multiple of KC.
  __device__ __forceinline__ void run(Smem& s, const View& A, const View& B, int32_t depth,
                                      int32_t (&acc)[TM][TN]) {
    const int nch = depth / KC;
    fetch(A, B, 0);
    stash(s, 0);
    __syncthreads();
    for (int ch = 0; ch < nch; ++ch) {
      const int buf = ch & 1;
      if (ch + 1 < nch) fetch(A, B, (ch + 1) * KC);
      if (read_block_flag<4>(s.flag[buf]))
        compute<false>(s, buf, acc);
      else
        compute<true>(s, buf, acc);
      if (ch + 1 < nch) stash(s, buf ^ 1);
      __syncthreads();
    }
  }

  __device__ __forceinline__ static bool full(const View& C) {
    return C.rlim >= BM && C.clim >= BN;
  }

  __device__ __forceinline__ static void load_c(const View& C, int32_t (&acc)[TM][TN]) {
    const int tx = threadIdx.x & 15;
    const int ty = threadIdx.x >> 4;
    if (full(C)) {
      const int32_t oc = ty * C.ld + tx;
#pragma unroll
      for (int r = 0; r < TM; ++r) {
        const int32_t* pr = C.p + (16 * r) * C.ld;
#pragma unroll
        for (int c = 0; c < TN; ++c) acc[r][c] = pr[oc + 16 * c];
      }
    } else {
#pragma unroll
      for (int r = 0; r < TM; ++r)
#pragma unroll
        for (int c = 0; c < TN; ++c)
          acc[r][c] = C.in(ty + 16 * r, tx + 16 * c) ? *C.at(ty + 16 * r, tx + 16 * c) : kInf;
    }
  }

  __device__ __forceinline__ static void store_c(const View& C, const int32_t (&acc)[TM][TN]) {
    const int tx = threadIdx.x & 15;
    const int ty = threadIdx.x >> 4;
    if (full(C)) {
      const int32_t oc = ty * C.ld + tx;
#pragma unroll
      for (int r = 0; r < TM; ++r) {
        int32_t* pr = C.p + (16 * r) * C.ld;
#pragma unroll
        for (int c = 0; c < TN; ++c) pr[oc + 16 * c] = acc[r][c];
      }
    } else {
#pragma unroll
      for (int r = 0; r < TM; ++r)
#pragma unroll
        for (int c = 0; c < TN; ++c)
          if (C.in(ty + 16 * r, tx + 16 * c)) *C.at(ty + 16 * r, tx + 16 * c) = acc[r][c];
    }
  }
};

// Close a T x T diagonal tile in place (Floyd-Warshall inside the tile), one
// 256-thread block.  Thread (tx, ty) keeps cells (ty+16r, tx+16q) in VGPRs.
// Two k steps per barrier: the owners of rows k, k+1 and columns k, k+1
// publish them to a double-buffered LDS line; every thread then forms
//   col'[i] = min(D[i][k+1], D[i][k] + D[k][k+1])   (column k+1 after step k)
//   row'[j] = min(D[k+1][j], D[k+1][k] + D[k][j])   (row k+1 after step k)
//   D[i][j] = min3(D[i][j], D[i][k] + D[k][j], col'[i] + row'[j])
// which is exactly steps k and k+1 of the sequential loop.  Line buffer p&1
// is rewritten at pair p+2, after every thread passed barrier p+1, so one
// barrier per pair suffices.  Steps with k >= n (padding) are skipped, so
// padded vertices are never intermediates.
template <int T>
__device__ __forceinline__ void close_tile(const View V) {
  constexpr int TM = T / 16;
  __shared__ int32_t rows[2][2][T];  // [buf][k, k+1][j]
  __shared__ int32_t cols[2][2][T];  // [buf][k, k+1][i]
  const int tx = threadIdx.x & 15;
  const int ty = threadIdx.x >> 4;
  int32_t acc[TM][TM];
#pragma unroll
  for (int r = 0; r < TM; ++r)
#pragma unroll
    for (int q = 0; q < TM; ++q)
      acc[r][q] = V.in(ty + 16 * r, tx + 16 * q) ? *V.at(ty + 16 * r, tx + 16 * q) : kInf;
  const int kmax = min(T, V.rlim);
#pragma unroll
  for (int kk = 0; kk < TM; ++kk) {
    for (int k16 = 0; k16 < 16; k16 += 2) {
      const int k = kk * 16 + k16;
      if (k >= kmax) break;
      const int b = (k >> 1) & 1;
      // rows k and k+1 live in the same register row kk (k16 even)
      if (ty == k16 || ty == k16 + 1) {
        const int h = ty - k16;
#pragma unroll
        for (int q = 0; q < TM; ++q) rows[b][h][tx + 16 * q] = acc[kk][q];
      }
      if (tx == k16 || tx == k16 + 1) {
        const int h = tx - k16;
#pragma unroll
        for (int r = 0; r < TM; ++r) cols[b][h][ty + 16 * r] = acc[r][kk];
      }
      __syncthreads();
      if (k + 1 < kmax) {
        const int32_t dkk1 = rows[b][0][k + 1];  // D[k][k+1]
        const int32_t dk1k = rows[b][1][k];      // D[k+1][k]
        int32_t rk[TM], rk1[TM], ck[TM], ck1[TM];
#pragma unroll
        for (int q = 0; q < TM; ++q) {
          rk[q] = rows[b][0][tx + 16 * q];
          rk1[q] = min(rows[b][1][tx + 16 * q], wadd(dk1k, rk[q]));
        }
#pragma unroll
        for (int r = 0; r < TM; ++r) {
          ck[r] = cols[b][0][ty + 16 * r];
          ck1[r] = min(cols[b][1][ty + 16 * r], wadd(ck[r], dkk1));
        }
#pragma unroll
        for (int r = 0; r < TM; ++r)
#pragma unroll
          for (int q = 0; q < TM; ++q) acc[r][q] = vmin3(acc[r][q], wadd(ck[r], rk[q]), wadd(ck1[r], rk1[q]));
      } else {
#pragma unroll
        for (int r = 0; r < TM; ++r) {
          const int32_t ckr = cols[b][0][ty + 16 * r];
#pragma unroll
          for (int q = 0; q < TM; ++q) acc[r][q] = min(acc[r][q], wadd(ckr, rows[b][0][tx + 16 * q]));
        }
      }
    }
  }
#pragma unroll
  for (int r = 0; r < TM; ++r)
#pragma unroll
    for (int q = 0; q < TM; ++q)
      if (V.in(ty + 16 * r, tx + 16 * q)) *V.at(ty + 16 * r, tx + 16 * q) = acc[r][q];
}

}  // namespace fwx
