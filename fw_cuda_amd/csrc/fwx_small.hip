// fwx_small.hip -- blocked Floyd-Warshall for small matrices: one launch per
// k-tile round on 32-wide k-tiles.
//
// Reference behaviour replaced: the three dependent launches per k-tile of
// fwTilingGPU (src/fw_tiling_gpu.cu:253-284) -- diagonal, row/column panels,
// independent tiles.  For small n the solve is a chain of N = n/32 rounds
// whose launches do little work each; the chain's cost is per-launch latency.
// Here round c is one launch.  A region block owns the R x R region of rows
// [I R, I R + R) and columns [J R, J R + R) (R = 64 or 128, a whole number
// of k-tiles) and, per round:
//
//   loads tile (c,c), its A panel X(I, c) (R x 32), its B panel X(c, J)
//   (32 x R) and its region X(I, J), all at once;
//   takes D*, the closure of tile (c,c) (the round's 32 k steps);
//   P_A = min(X(I,c), X(I,c) (x) D*),  P_B = min(X(c,J), D* (x) X(c,J));
//   stores min(X(I,J), P_A (x) P_B)
//
// with every product's operands in LDS.  The closure is the round's longest
// dependent chain (32 steps), so it is taken off the chain of launches: one
// more block per launch, the look-ahead block, computes round c's result on
// tile (c+1, c+1) -- from the two panel tiles P(c+1,c) and P(c,c+1), which it
// computes itself -- closes it and stores D* of round c+1 there; the region
// block holding that tile does not store it.  Tile (c,c) therefore holds D*
// when round c's launch starts (tile (0,0) is closed by a one-block launch
// before round 0), and region blocks only read it.  For a 32-tile (i,j) of the region
// this is the three-phase round: a cross tile gets min(X, P(i,c) (x) P(c,j));
// a panel tile gets min(P(c,j), D* (x) P(c,j)) = P(c,j) (and likewise
// P(i,c)), since D* (x) D* >= D* (triangle inequality inside the closed
// tile); tile (c,c) gets min(X(c,c), D* (x) D*) = D* (each entry of D* is
// a direct entry of X(c,c) or splits at an intermediate vertex of the tile).
//
// Exactness with the in-place races.  Tile (c,c) and the panel tiles are read
// by many blocks while the one block whose region holds them overwrites them.
// A read of such a cell returns the round's input value or its output value
// (aligned 4-byte cells are never torn), so a block sees, per tile, some mix
// Y with out <= Y <= in cellwise.  With no negative cycle:
//   * closing a mix Y of X(c,c) and D* gives D* (closure is monotone and
//     closes both ends to D*);
//   * for a panel, P = min(X, X (x) D*) and P (x) D* >= X (x) D* >= P, so
//     min(Y, Y (x) D*) lies between min(P, P (x) D*) = P and
//     min(X, X (x) D*) = P;
// so every block computes exactly the panels of the three-phase order,
// whichever values it read.  A region is written by its own block only,
// after it read it, except tile (c+1, c+1), which only the look-ahead block
// writes (with D* of round c+1: at most the round's value there, and what
// round c+1 would compute from it).  Tile (c, c) already holds D*, and the
// region block that stores it writes min(D*, D* (x) D*) = D*, the same
// values.  Rounds are ordered by the launch boundaries.
//
// Relaxation form per block and round: the block ORs every operand it loaded
// and takes the pk64 form when all lie in [0, 2^30), else min3.  Every value
// it derives (closure, panels, region) is a min over its operands and their
// sums, so it stays in that range too; the choice is exact for what the
// block computes, and no range pass over the matrix or host round trip is
// needed.  Ragged n: operand cells outside the matrix are INF, k steps >= n
// are skipped, and tiles written to LDS from registers (D*, the panels) are
// masked to INF outside the matrix, so values that grew in padding rows or
// columns never reach a real cell.
#include <hip/hip_runtime.h>

#include <cstdlib>

#include "fwx_device.hpp"
#include "fwx_internal.hpp"
#include "fwx_tile.hpp"
#include "fwx_tile16.hpp"

namespace fwx {

namespace {

constexpr int kT = 32;        // k-tile
constexpr int kLda = kT + 4;  // A rows in LDS: 16-byte aligned groups of 4 k

template <int R>
struct SmallSmem {
  alignas(16) int32_t dA[kT * kLda];  // D*, row-major (A operand)
  alignas(16) int2 dB[(kT / 2) * kT]; // D*, k pairs (B operand)
  alignas(16) int32_t pA[R * kLda];   // A panel: X(I,c), then P_A
  alignas(16) int2 pB[(kT / 2) * R];  // B panel: X(c,J), then P_B
  int32_t lines[8 * kT];              // closure
  uint32_t wflags[16];                // form vote, one word per wave
};

// Block-wide vote on the operands a block loaded, with one barrier: bit 0 =
// some thread holds a value outside [0, 2^30) (no pk64), bit 1 = some thread
// holds one >= kSmall16 (no pk16).
template <int NT>
__device__ __forceinline__ uint32_t form_vote(uint32_t* wflags, uint32_t o);

// Thread (tx, ty) = (tid % 16, tid / 16) of NT threads holds cells
// (ty + (NT/16) r, tx + 16 q) of a BM x BN window -- for the 32 x 32 tile
// (c,c) the layout of close_regs<32, NT>.
template <int NT, int BM, int BN>
using Cells = int32_t[BM / (NT / 16)][BN / 16];

template <int NT, int BM, int BN>
__device__ __forceinline__ void load_cells(const int32_t* D, int32_t n, int32_t r0, int32_t c0, int tx, int ty,
                                           Cells<NT, BM, BN>& t) {
  constexpr int RS = NT / 16;
#pragma unroll
  for (int r = 0; r < BM / RS; ++r)
#pragma unroll
    for (int q = 0; q < BN / 16; ++q) {
      const int32_t row = r0 + ty + RS * r, col = c0 + tx + 16 * q;
      t[r][q] = (row < n && col < n) ? D[(int64_t)row * n + col] : kInf;
    }
}

// Cells of the 32 x 32 diagonal tile starting at (skip, skip) are not
// stored (skip < 0: none).
template <int NT, int BM, int BN>
__device__ __forceinline__ void store_cells(int32_t* D, int32_t n, int32_t r0, int32_t c0, int tx, int ty,
                                            const Cells<NT, BM, BN>& t, int32_t skip = -kT) {
  constexpr int RS = NT / 16;
#pragma unroll
  for (int r = 0; r < BM / RS; ++r)
#pragma unroll
    for (int q = 0; q < BN / 16; ++q) {
      const int32_t row = r0 + ty + RS * r, col = c0 + tx + 16 * q;
      const bool skipped = (uint32_t)(row - skip) < (uint32_t)kT && (uint32_t)(col - skip) < (uint32_t)kT;
      if (row < n && col < n && !skipped) D[(int64_t)row * n + col] = t[r][q];
    }
}

template <int NT, int BM, int BN>
__device__ __forceinline__ uint32_t or_cells(const Cells<NT, BM, BN>& t) {
  uint32_t o = 0;
#pragma unroll
  for (int r = 0; r < BM / (NT / 16); ++r)
#pragma unroll
    for (int q = 0; q < BN / 16; ++q) o |= (uint32_t)t[r][q];
  return o;
}

// Registers -> LDS as an A operand (row-major, BM x 32) or a B operand
// (k pairs, 32 x BN); cells at window row >= rl or column >= cl as INF.
template <int NT, int BM>
__device__ __forceinline__ void to_a(int32_t* A, const Cells<NT, BM, kT>& t, int tx, int ty, int32_t rl,
                                     int32_t cl) {
  constexpr int RS = NT / 16;
#pragma unroll
  for (int r = 0; r < BM / RS; ++r)
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int row = ty + RS * r, col = tx + 16 * q;
      A[row * kLda + col] = (row < rl && col < cl) ? t[r][q] : kInf;
    }
}

template <int NT, int BN>
__device__ __forceinline__ void to_b(int2* B, const Cells<NT, kT, BN>& t, int tx, int ty, int32_t rl, int32_t cl) {
  constexpr int RS = NT / 16;
  int32_t* b = reinterpret_cast<int32_t*>(B);
#pragma unroll
  for (int r = 0; r < kT / RS; ++r)
#pragma unroll
    for (int q = 0; q < BN / 16; ++q) {
      const int row = ty + RS * r, col = tx + 16 * q;
      b[((row >> 1) * BN + col) * 2 + (row & 1)] = (row < rl && col < cl) ? t[r][q] : kInf;
    }
}

// acc <- min(acc, A (x) B) over the 32 k of the round; A is BM x 32
// row-major, B 32 x BN in k pairs, both in LDS.  Eight groups of 4 k: each
// group's operands (one 16-byte A read per row, two 8-byte B reads per
// column) are loaded while the previous group is computed, so the LDS
// latency overlaps the relaxations instead of stalling the wave.
template <int NT, int BM, int BN, bool PK>
__device__ __forceinline__ void prod_lds(const int32_t* A, const int2* B, Cells<NT, BM, BN>& acc, int tx, int ty) {
  constexpr int RS = NT / 16, TR = BM / RS, TQ = BN / 16;
  int4 a[2][TR];
  int2 b[2][2][TQ];
  auto load = [&](int buf, int g) {
#pragma unroll
    for (int r = 0; r < TR; ++r) a[buf][r] = *reinterpret_cast<const int4*>(&A[(ty + RS * r) * kLda + 4 * g]);
#pragma unroll
    for (int q = 0; q < TQ; ++q) {
      b[buf][0][q] = B[(2 * g) * BN + tx + 16 * q];
      b[buf][1][q] = B[(2 * g + 1) * BN + tx + 16 * q];
    }
  };
  load(0, 0);
#pragma unroll 2
  for (int g = 0; g < kT / 4; ++g) {
    const int cur = g & 1;
    if (g + 1 < kT / 4) load(cur ^ 1, g + 1);
#pragma unroll
    for (int r = 0; r < TR; ++r) {
      const uint64_t a0 = pack2(a[cur][r].x, a[cur][r].y), a1 = pack2(a[cur][r].z, a[cur][r].w);
#pragma unroll
      for (int q = 0; q < TQ; ++q) {
        const uint64_t b0 = pack2(b[cur][0][q].x, b[cur][0][q].y), b1 = pack2(b[cur][1][q].x, b[cur][1][q].y);
        acc[r][q] = PK ? relax2_pk64(acc[r][q], a0, b0) : relax2_min3(acc[r][q], a0, b0);
        acc[r][q] = PK ? relax2_pk64(acc[r][q], a1, b1) : relax2_min3(acc[r][q], a1, b1);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// The pk16 form (fwx_tile16.hpp) for blocks whose loaded operands are all
// below kSmall16: every sum of two values the block derives is then below
// 2 kSmall16 <= the pk16 cap (each derived value is a min over loaded values
// and such sums, and never exceeds the value it replaces), so nothing
// saturates and no check is needed afterwards.  Such a block holds no cell
// outside the matrix (those load as INF), so its staging needs no mask.
// A operand: (v + BIAS) in both 16-bit lanes, the int32 A layout.  B operand:
// per k pair, column pairs (j, j + 16) with j = (cp % 16) + 32 (cp / 16), one
// 8-byte slot each (low word row 2p, high word row 2p + 1) -- the pairs a
// thread's cells form (columns tx + 16 q and tx + 16 (q + 1), q even).
constexpr uint32_t kSmall16 = 0x1000u;

template <int NT, int BM>
__device__ __forceinline__ void to_a16(uint32_t* A, const Cells<NT, BM, kT>& t, int tx, int ty) {
  constexpr int RS = NT / 16;
#pragma unroll
  for (int r = 0; r < BM / RS; ++r)
#pragma unroll
    for (int q = 0; q < 2; ++q) A[(ty + RS * r) * kLda + tx + 16 * q] = ((uint32_t)t[r][q] + kP16Bias) * 0x10001u;
}

// B slot (p, cp) low or high word: 16 lanes of consecutive tx write
// consecutive slots, the next 16 (row + 1) the other words of the same slots.
template <int NT, int BN>
__device__ __forceinline__ void to_b16(uint64_t* B, const Cells<NT, kT, BN>& t, int tx, int ty) {
  constexpr int RS = NT / 16;
  uint32_t* b = reinterpret_cast<uint32_t*>(B);
#pragma unroll
  for (int r = 0; r < kT / RS; ++r)
#pragma unroll
    for (int qp = 0; qp < BN / 32; ++qp) {
      const int row = ty + RS * r;
      b[((row >> 1) * (BN / 2) + tx + 16 * qp) * 2 + (row & 1)] =
          (uint32_t)t[r][2 * qp] | ((uint32_t)t[r][2 * qp + 1] << 16);
    }
}

template <int NT, int BM, int BN>
using Pairs = uint32_t[BM / (NT / 16)][BN / 32];

template <int NT, int BM, int BN>
__device__ __forceinline__ void enc16(const Cells<NT, BM, BN>& t, Pairs<NT, BM, BN>& c) {
#pragma unroll
  for (int r = 0; r < BM / (NT / 16); ++r)
#pragma unroll
    for (int qp = 0; qp < BN / 32; ++qp)
      c[r][qp] = ((uint32_t)t[r][2 * qp] | ((uint32_t)t[r][2 * qp + 1] << 16)) + kP16Bias2;
}

template <int NT, int BM, int BN>
__device__ __forceinline__ void dec16(const Pairs<NT, BM, BN>& c, Cells<NT, BM, BN>& t) {
#pragma unroll
  for (int r = 0; r < BM / (NT / 16); ++r)
#pragma unroll
    for (int qp = 0; qp < BN / 32; ++qp) {
      const uint32_t v = c[r][qp] - kP16Bias2;
      t[r][2 * qp] = (int32_t)(v & 0xFFFFu);
      t[r][2 * qp + 1] = (int32_t)(v >> 16);
    }
}

// acc <- min(acc, A (x) B) over the round's 32 k in the pk16 form; the next
// group of 4 k is read while the current one is computed.
template <int NT, int BM, int BN>
__device__ __forceinline__ void prod_lds16(const uint32_t* A, const uint64_t* B, Pairs<NT, BM, BN>& acc, int tx,
                                           int ty) {
  constexpr int RS = NT / 16, TR = BM / RS, TP = BN / 32;
  uint4 a[2][TR];
  uint64_t b[2][2][TP];
  auto load = [&](int buf, int g) {
#pragma unroll
    for (int r = 0; r < TR; ++r) a[buf][r] = *reinterpret_cast<const uint4*>(&A[(ty + RS * r) * kLda + 4 * g]);
#pragma unroll
    for (int qp = 0; qp < TP; ++qp) {
      b[buf][0][qp] = B[(2 * g) * (BN / 2) + tx + 16 * qp];
      b[buf][1][qp] = B[(2 * g + 1) * (BN / 2) + tx + 16 * qp];
    }
  };
  load(0, 0);
#pragma unroll 2
  for (int g = 0; g < kT / 4; ++g) {
    const int cur = g & 1;
    if (g + 1 < kT / 4) load(cur ^ 1, g + 1);
#pragma unroll
    for (int r = 0; r < TR; ++r) {
      const uint64_t a01 = (uint64_t)a[cur][r].x | ((uint64_t)a[cur][r].y << 32);
      const uint64_t a23 = (uint64_t)a[cur][r].z | ((uint64_t)a[cur][r].w << 32);
#pragma unroll
      for (int qp = 0; qp < TP; ++qp) {
        acc[r][qp] = relax4_pk16(acc[r][qp], a01, b[cur][0][qp]);
        acc[r][qp] = relax4_pk16(acc[r][qp], a23, b[cur][1][qp]);
      }
    }
  }
}

template <int NT, int BM, int BN>
__device__ __forceinline__ void prod16(const uint32_t* A, const uint64_t* B, Cells<NT, BM, BN>& t, int tx, int ty) {
  Pairs<NT, BM, BN> acc;
  enc16<NT, BM, BN>(t, acc);
  prod_lds16<NT, BM, BN>(A, B, acc, tx, ty);
  dec16<NT, BM, BN>(acc, t);
}

template <int NT>
__device__ __forceinline__ uint32_t form_vote(uint32_t* wflags, uint32_t o) {
  const uint32_t f = (__ballot((o & kPk64Mask) != 0u) ? 1u : 0u) | (__ballot(o >= kSmall16) ? 2u : 0u);
  if ((threadIdx.x & 63) == 0) wflags[threadIdx.x >> 6] = f;
  __syncthreads();
  uint32_t all = 0;
#pragma unroll
  for (int w = 0; w < NT / 64; ++w) all |= wflags[w];
  return all;
}

// Per round: the two panel products run at once, one per half of the block
// (NT/2 threads each, wave-uniform), then the region product on all NT.
template <int R, int NT, bool PK>
__device__ __forceinline__ void small_products(SmallSmem<R>& s, Cells<NT / 2, R, kT>& pa, Cells<NT / 2, kT, R>& pb,
                                               Cells<NT, R, R>& xc, bool upper, int32_t n, int32_t c0,
                                               int32_t r0, int32_t col0) {
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4, th = (threadIdx.x & (NT / 2 - 1)) >> 4;
  if (!upper)
    prod_lds<NT / 2, R, kT, PK>(s.pA, s.dB, pa, tx, th);  // P_A = min(X(I,c), X(I,c) (x) D*)
  else
    prod_lds<NT / 2, kT, R, PK>(s.dA, s.pB, pb, tx, th);  // P_B = min(X(c,J), D* (x) X(c,J))
  __syncthreads();
  if (!upper)
    to_a<NT / 2, R>(s.pA, pa, tx, th, n - r0, n - c0);
  else
    to_b<NT / 2, R>(s.pB, pb, tx, th, n - c0, n - col0);
  __syncthreads();
  prod_lds<NT, R, R, PK>(s.pA, s.pB, xc, tx, ty);
}

// The look-ahead block of round c (when c + 1 < N): the round's result on
// tile (c+1, c+1) is min(X(c+1,c+1), P(c+1,c) (x) P(c,c+1)) with the round's
// two panel tiles; the block computes those itself (one per half), then the
// tile, closes it -- the next round's D* -- and stores it.  The region that
// holds tile (c+1, c+1) does not store it this round.
template <int R, int NT, bool PK>
__device__ __forceinline__ void lookahead_products(SmallSmem<R>& s, Cells<NT / 2, kT, kT>& p,
                                                   Cells<NT, kT, kT>& y, bool upper, int32_t n, int32_t c0,
                                                   int32_t e0) {
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4, th = (threadIdx.x & (NT / 2 - 1)) >> 4;
  if (!upper)
    prod_lds<NT / 2, kT, kT, PK>(s.pA, s.dB, p, tx, th);  // P(c+1,c)
  else
    prod_lds<NT / 2, kT, kT, PK>(s.dA, s.pB, p, tx, th);  // P(c,c+1)
  __syncthreads();
  if (!upper)
    to_a<NT / 2, kT>(s.pA, p, tx, th, n - e0, n - c0);
  else
    to_b<NT / 2, kT>(s.pB, p, tx, th, n - c0, n - e0);
  __syncthreads();
  prod_lds<NT, kT, kT, PK>(s.pA, s.pB, y, tx, ty);
}

// small_products in the pk16 form (operands already staged as pk16).
template <int R, int NT>
__device__ __forceinline__ void small_products16(SmallSmem<R>& s, Cells<NT / 2, R, kT>& pa, Cells<NT / 2, kT, R>& pb,
                                                 Cells<NT, R, R>& xc, bool upper) {
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4, th = (threadIdx.x & (NT / 2 - 1)) >> 4;
  uint32_t* dA = reinterpret_cast<uint32_t*>(s.dA);
  uint64_t* dB = reinterpret_cast<uint64_t*>(s.dB);
  uint32_t* pA = reinterpret_cast<uint32_t*>(s.pA);
  uint64_t* pB = reinterpret_cast<uint64_t*>(s.pB);
  if (!upper)
    prod16<NT / 2, R, kT>(pA, dB, pa, tx, th);  // P_A = min(X(I,c), X(I,c) (x) D*)
  else
    prod16<NT / 2, kT, R>(dA, pB, pb, tx, th);  // P_B = min(X(c,J), D* (x) X(c,J))
  __syncthreads();
  if (!upper)
    to_a16<NT / 2, R>(pA, pa, tx, th);
  else
    to_b16<NT / 2, R>(pB, pb, tx, th);
  __syncthreads();
  prod16<NT, R, R>(pA, pB, xc, tx, ty);
}

// lookahead_products in the pk16 form.
template <int R, int NT>
__device__ __forceinline__ void lookahead_products16(SmallSmem<R>& s, Cells<NT / 2, kT, kT>& p,
                                                     Cells<NT, kT, kT>& y, bool upper) {
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4, th = (threadIdx.x & (NT / 2 - 1)) >> 4;
  uint32_t* dA = reinterpret_cast<uint32_t*>(s.dA);
  uint64_t* dB = reinterpret_cast<uint64_t*>(s.dB);
  uint32_t* pA = reinterpret_cast<uint32_t*>(s.pA);
  uint64_t* pB = reinterpret_cast<uint64_t*>(s.pB);
  if (!upper)
    prod16<NT / 2, kT, kT>(pA, dB, p, tx, th);  // P(c+1,c)
  else
    prod16<NT / 2, kT, kT>(dA, pB, p, tx, th);  // P(c,c+1)
  __syncthreads();
  if (!upper)
    to_a16<NT / 2, kT>(pA, p, tx, th);
  else
    to_b16<NT / 2, kT>(pB, p, tx, th);
  __syncthreads();
  prod16<NT, kT, kT>(pA, pB, y, tx, ty);
}

template <int R, int NT>
__device__ __forceinline__ void lookahead_block(SmallSmem<R>& s, int32_t* D, int32_t n, int32_t c,
                                                bool small16_enabled) {
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4, th = (threadIdx.x & (NT / 2 - 1)) >> 4;
  const bool upper = threadIdx.x >= NT / 2;
  const int32_t c0 = c * kT, e0 = c0 + kT;
  Cells<NT, kT, kT> d, y;
  Cells<NT / 2, kT, kT> p;
  load_cells<NT, kT, kT>(D, n, c0, c0, tx, ty, d);
  if (!upper)
    load_cells<NT / 2, kT, kT>(D, n, e0, c0, tx, th, p);
  else
    load_cells<NT / 2, kT, kT>(D, n, c0, e0, tx, th, p);
  load_cells<NT, kT, kT>(D, n, e0, e0, tx, ty, y);
  const uint32_t o = or_cells<NT, kT, kT>(d) | or_cells<NT / 2, kT, kT>(p) | or_cells<NT, kT, kT>(y);
  const uint32_t vote = form_vote<NT>(s.wflags, o);
  const bool pk = !(vote & 1u);
  const bool p16 = small16_enabled && !(vote & 2u);
  if (p16) {
    to_a16<NT, kT>(reinterpret_cast<uint32_t*>(s.dA), d, tx, ty);
    to_b16<NT, kT>(reinterpret_cast<uint64_t*>(s.dB), d, tx, ty);
    if (!upper)
      to_a16<NT / 2, kT>(reinterpret_cast<uint32_t*>(s.pA), p, tx, th);
    else
      to_b16<NT / 2, kT>(reinterpret_cast<uint64_t*>(s.pB), p, tx, th);
    __syncthreads();
    lookahead_products16<R, NT>(s, p, y, upper);
    close_regs<kT, NT>(y, s.lines, min(kT, n - e0));
    store_cells<NT, kT, kT>(D, n, e0, e0, tx, ty, y);
    return;
  }
  to_a<NT, kT>(s.dA, d, tx, ty, n - c0, n - c0);
  to_b<NT, kT>(s.dB, d, tx, ty, n - c0, n - c0);
  if (!upper)
    to_a<NT / 2, kT>(s.pA, p, tx, th, kT, kT);
  else
    to_b<NT / 2, kT>(s.pB, p, tx, th, kT, kT);
  __syncthreads();
  if (pk)
    lookahead_products<R, NT, true>(s, p, y, upper, n, c0, e0);
  else
    lookahead_products<R, NT, false>(s, p, y, upper, n, c0, e0);
  close_regs<kT, NT>(y, s.lines, min(kT, n - e0));
  store_cells<NT, kT, kT>(D, n, e0, e0, tx, ty, y);
}

// Round c: block 0 is the look-ahead block (when c + 1 < N), block b > 0 the
// region b - 1 of M x M.  Tile (c, c) holds D* already (the previous
// round's look-ahead block stored it; k_small_close0 for round 0).  Thread
// (tx, ty) holds the region cells and the cells of tile (c,c) in the
// NT-thread layout; the lower half of the block (th = its row within the
// half) holds the A panel, the upper half the B panel.  The operands are
// loaded at once, then one block-wide vote picks the form.
template <int R, int NT>
__global__ __launch_bounds__(NT) void k_small_round(int32_t* __restrict__ D, int32_t n, int32_t c, int32_t M,
                                                    int32_t N, int32_t small16_enabled) {
  __shared__ SmallSmem<R> s;
  const bool ahead = c + 1 < N;
  if (ahead && blockIdx.x == 0) {
    lookahead_block<R, NT>(s, D, n, c, small16_enabled != 0);
    return;
  }
  const int b = blockIdx.x - (ahead ? 1 : 0);
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4, th = (threadIdx.x & (NT / 2 - 1)) >> 4;
  const bool upper = threadIdx.x >= NT / 2;
  const int I = b / M, J = b - (b / M) * M;
  const int32_t c0 = c * kT, r0 = I * R, col0 = J * R;
  Cells<NT, kT, kT> d;
  Cells<NT / 2, R, kT> pa;
  Cells<NT / 2, kT, R> pb;
  Cells<NT, R, R> xc;
  load_cells<NT, kT, kT>(D, n, c0, c0, tx, ty, d);
  if (!upper)
    load_cells<NT / 2, R, kT>(D, n, r0, c0, tx, th, pa);
  else
    load_cells<NT / 2, kT, R>(D, n, c0, col0, tx, th, pb);
  load_cells<NT, R, R>(D, n, r0, col0, tx, ty, xc);
  const uint32_t o = or_cells<NT, kT, kT>(d) | (upper ? or_cells<NT / 2, kT, R>(pb) : or_cells<NT / 2, R, kT>(pa)) |
                     or_cells<NT, R, R>(xc);
  const uint32_t vote = form_vote<NT>(s.wflags, o);
  const bool pk = !(vote & 1u);
  const bool p16 = small16_enabled && !(vote & 2u);
  if (p16) {
    to_a16<NT, kT>(reinterpret_cast<uint32_t*>(s.dA), d, tx, ty);
    to_b16<NT, kT>(reinterpret_cast<uint64_t*>(s.dB), d, tx, ty);
    if (!upper)
      to_a16<NT / 2, R>(reinterpret_cast<uint32_t*>(s.pA), pa, tx, th);
    else
      to_b16<NT / 2, R>(reinterpret_cast<uint64_t*>(s.pB), pb, tx, th);
    __syncthreads();
    small_products16<R, NT>(s, pa, pb, xc, upper);
    store_cells<NT, R, R>(D, n, r0, col0, tx, ty, xc, ahead ? c0 + kT : -kT);
    return;
  }
  to_a<NT, kT>(s.dA, d, tx, ty, n - c0, n - c0);
  to_b<NT, kT>(s.dB, d, tx, ty, n - c0, n - c0);
  if (!upper)
    to_a<NT / 2, R>(s.pA, pa, tx, th, R, kT);
  else
    to_b<NT / 2, R>(s.pB, pb, tx, th, kT, R);
  __syncthreads();
  if (pk)
    small_products<R, NT, true>(s, pa, pb, xc, upper, n, c0, r0, col0);
  else
    small_products<R, NT, false>(s, pa, pb, xc, upper, n, c0, r0, col0);
  // tile (c+1, c+1) is the look-ahead block's to store
  store_cells<NT, R, R>(D, n, r0, col0, tx, ty, xc, ahead ? c0 + kT : -kT);
}

// Tile (0, 0) closed once before round 0, so that every round's region
// blocks only read a closed diagonal tile.
__global__ __launch_bounds__(256) void k_small_close0(int32_t* __restrict__ D, int32_t n) {
  close_tile<kT, 256>(make_view(D, n, 0, 0));
}

// Region width R (FWX_SMALL_R = 64 or 128; default 64: measured faster at
// every n from 256 to 4096) and threads per block (FWX_SMALL_NT = 256 or
// 512; default 512).
int small_region(int32_t) {
  const char* e = std::getenv("FWX_SMALL_R");
  const int v = e ? std::atoi(e) : 0;
  return v == 128 ? 128 : 64;
}

int small_threads() {
  const char* e = std::getenv("FWX_SMALL_NT");
  const int v = e ? std::atoi(e) : 0;
  return v == 256 ? 256 : 512;
}

// FWX_SMALL16=0 keeps every block on pk64 / min3 (read at every call).
int small16() {
  const char* e = std::getenv("FWX_SMALL16");
  return (e && e[0] == '0') ? 0 : 1;
}

}  // namespace

int enqueue_small_rounds(int32_t* d, int32_t n, hipStream_t st) {
  const int N = (n + kT - 1) / kT;
  const int R = small_region(n), NT = small_threads(), p16 = small16();
  const int M = (n + R - 1) / R;
  k_small_close0<<<1, 256, 0, st>>>(d, n);
  FWX_LAUNCHED("k_small_close0");
  for (int c = 0; c < N; ++c) {
    const int blocks = M * M + (c + 1 < N ? 1 : 0);
    if (R == 64 && NT == 256)
      k_small_round<64, 256><<<blocks, 256, 0, st>>>(d, n, c, M, N, p16);
    else if (R == 64)
      k_small_round<64, 512><<<blocks, 512, 0, st>>>(d, n, c, M, N, p16);
    else if (NT == 256)
      k_small_round<128, 256><<<blocks, 256, 0, st>>>(d, n, c, M, N, p16);
    else
      k_small_round<128, 512><<<blocks, 512, 0, st>>>(d, n, c, M, N, p16);
    FWX_LAUNCHED("k_small_round");
  }
  return FWX_OK;
}

int small_region_knob(int32_t n) { return small_region(n) | (small_threads() << 8) | (small16() << 16); }

}  // namespace fwx
