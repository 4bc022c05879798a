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
  uint32_t wflags[16];                // votes: one word per wave, two sets
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

// prod_lds16 with every operand of the 32 k read before the first
// relaxation: one LDS latency per product instead of one per group of 4 k,
// for the small operand sets of the 32 x 32 closure (TR = TP = 1 at 512
// threads: 64 VGPRs of operands).
template <int NT, int BM, int BN>
__device__ __forceinline__ void prod_lds16_all(const uint32_t* A, const uint64_t* B, Pairs<NT, BM, BN>& acc, int tx,
                                               int ty) {
  constexpr int RS = NT / 16, TR = BM / RS, TP = BN / 32;
  static_assert(TR * TP == 1, "operands of one cell pair per thread");
  uint4 a[kT / 4];
  uint64_t b[kT / 2];
#pragma unroll
  for (int g = 0; g < kT / 4; ++g) a[g] = *reinterpret_cast<const uint4*>(&A[ty * kLda + 4 * g]);
#pragma unroll
  for (int p = 0; p < kT / 2; ++p) b[p] = B[p * (BN / 2) + tx];
#pragma unroll
  for (int g = 0; g < kT / 4; ++g) {
    const uint64_t a01 = (uint64_t)a[g].x | ((uint64_t)a[g].y << 32);
    const uint64_t a23 = (uint64_t)a[g].z | ((uint64_t)a[g].w << 32);
    acc[0][0] = relax4_pk16(acc[0][0], a01, b[2 * g]);
    acc[0][0] = relax4_pk16(acc[0][0], a23, b[2 * g + 1]);
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

// Closure of the 32 x 32 tile y (all cells below kSmall16) by min-plus
// squaring in the pk16 form: y <- min(y, y (x) y) until a squaring changes
// no cell, at most five times.  After s squarings y holds, per cell, the
// shortest walk of at most 2^s edges over the tile's vertices; a squaring
// that changes nothing means y obeys every triangle inequality, i.e. it is
// the closure; after five, walks of 32 edges cover every simple path of the
// 32 vertices.  So with no negative cycle the result is the closure the
// 32 k steps of Floyd-Warshall give, in 2-5 dependent products instead of 32
// dependent steps.  Sums stay below 2 kSmall16 (every value is a min over
// the loaded ones and such sums), so nothing saturates.
// One barrier per squaring: each squaring stages its result as the next
// one's operands (into the other of the D* and panel staging buffers) before
// the vote, whose barrier then also publishes them.  Buffer b is rewritten
// two squarings after it was read, with a vote barrier in between; the
// first squaring's buffers (D*, then the panels) were last read before
// barriers of the look-ahead products.
template <int R, int NT>
__device__ __forceinline__ void close_sq16(SmallSmem<R>& s, Cells<NT, kT, kT>& y) {
  static_assert(2 * (NT / 64) <= 16, "two sets of vote words");
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  constexpr int TR = kT / (NT / 16), TP = kT / 32;
  Pairs<NT, kT, kT> acc, was;
  enc16<NT, kT, kT>(y, acc);
  to_a16<NT, kT>(reinterpret_cast<uint32_t*>(s.dA), y, tx, ty);
  to_b16<NT, kT>(reinterpret_cast<uint64_t*>(s.dB), y, tx, ty);
  __syncthreads();
  for (int it = 0; it < 5; ++it) {
    const uint32_t* A = reinterpret_cast<const uint32_t*>((it & 1) ? s.pA : s.dA);
    const uint64_t* B = reinterpret_cast<const uint64_t*>((it & 1) ? s.pB : s.dB);
#pragma unroll
    for (int r = 0; r < TR; ++r)
#pragma unroll
      for (int p = 0; p < TP; ++p) was[r][p] = acc[r][p];
    if constexpr (NT == 512)
      prod_lds16_all<NT, kT, kT>(A, B, acc, tx, ty);
    else
      prod_lds16<NT, kT, kT>(A, B, acc, tx, ty);
    uint32_t changed = 0;
#pragma unroll
    for (int r = 0; r < TR; ++r)
#pragma unroll
      for (int p = 0; p < TP; ++p) changed |= acc[r][p] ^ was[r][p];
    dec16<NT, kT, kT>(acc, y);
    if (it + 1 < 5) {
      to_a16<NT, kT>(reinterpret_cast<uint32_t*>((it & 1) ? s.dA : s.pA), y, tx, ty);
      to_b16<NT, kT>(reinterpret_cast<uint64_t*>((it & 1) ? s.dB : s.pB), y, tx, ty);
    }
    // the two halves of the vote words alternate, so a wave's next vote
    // cannot overwrite a word another wave has still to read
    if (!form_vote<NT>(s.wflags + (it & 1) * (NT / 64), changed ? kSmall16 : 0u)) break;
  }
}

// Read back, in the NT/2-thread layout of one half, a panel staged from the
// NT-thread layout by to_a (BM x 32, unmasked) or to_b (32 x BN, unmasked).
template <int NT2, int BM>
__device__ __forceinline__ void from_a(const int32_t* A, Cells<NT2, BM, kT>& t, int tx, int th) {
  constexpr int RS = NT2 / 16;
#pragma unroll
  for (int r = 0; r < BM / RS; ++r)
#pragma unroll
    for (int q = 0; q < 2; ++q) t[r][q] = A[(th + RS * r) * kLda + tx + 16 * q];
}

template <int NT2, int BN>
__device__ __forceinline__ void from_b(const int2* B, Cells<NT2, kT, BN>& t, int tx, int th) {
  constexpr int RS = NT2 / 16;
  const int32_t* b = reinterpret_cast<const int32_t*>(B);
#pragma unroll
  for (int r = 0; r < kT / RS; ++r)
#pragma unroll
    for (int q = 0; q < BN / 16; ++q) {
      const int row = th + RS * r, col = tx + 16 * q;
      t[r][q] = b[((row >> 1) * BN + col) * 2 + (row & 1)];
    }
}

// The round's product on a BM x BN window of rows [r0, r0 + BM) and columns
// [col0, col0 + BN) by the single-panel identity:
//   P_A = min(X(I,c), X(I,c) (x) D*)     (all NT threads)
//   C  <- min(C, P_A (x) X(c,J))
// which equals min(C, P_A (x) P_B) of the three-phase order when no value
// is negative: P_A (x) P_B = min(P_A (x) X(c,J), P_A (x) D* (x) X(c,J)) and
// P_A (x) D* >= X(I,c) (x) D* >= P_A, because D* (x) D* >= D* in a closed
// tile.  The same bound makes a read of X(c,J) that the block holding it
// has overwritten with P_B give the same result: P_A (x) P_B <= P_A (x) Y
// <= P_A (x) X(c,J) for any mix Y.  Staged on entry (and published by a
// barrier): D* as a B operand in dB, X(I,c) as A in pA, X(c,J) as B in pB
// -- pk16 (FORM 2) or int32 pk64 (FORM 1); qa holds X(I,c), c the window.
template <int R, int NT, int BM, int BN, int FORM>
__device__ __forceinline__ void single_panel(SmallSmem<R>& s, Cells<NT, BM, kT>& qa, Cells<NT, BM, BN>& xc,
                                             int32_t n, int32_t c0, int32_t r0) {
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  if constexpr (FORM == 2) {
    uint32_t* pA = reinterpret_cast<uint32_t*>(s.pA);
    prod16<NT, BM, kT>(pA, reinterpret_cast<const uint64_t*>(s.dB), qa, tx, ty);
    __syncthreads();
    to_a16<NT, BM>(pA, qa, tx, ty);
    __syncthreads();
    prod16<NT, BM, BN>(pA, reinterpret_cast<const uint64_t*>(s.pB), xc, tx, ty);
  } else {
    prod_lds<NT, BM, kT, true>(s.pA, s.dB, qa, tx, ty);
    __syncthreads();
    to_a<NT, BM>(s.pA, qa, tx, ty, n - r0, n - c0);
    __syncthreads();
    prod_lds<NT, BM, BN, true>(s.pA, s.pB, xc, tx, ty);
  }
}

// Stage D*, X(I,c) and X(c,J) for single_panel (FORM 2 or 1) or for the
// two-panel min3 products (FORM 0, which also stages D* as an A operand);
// then one barrier.  D* is masked to INF outside the matrix; the loaded
// panels already are.
template <int R, int NT, int BM, int BN, int FORM>
__device__ __forceinline__ void stage_round(SmallSmem<R>& s, const Cells<NT, kT, kT>& d,
                                            const Cells<NT, BM, kT>& qa, const Cells<NT, kT, BN>& qb, int32_t n,
                                            int32_t c0) {
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  if constexpr (FORM == 2) {
    to_b16<NT, kT>(reinterpret_cast<uint64_t*>(s.dB), d, tx, ty);
    to_a16<NT, BM>(reinterpret_cast<uint32_t*>(s.pA), qa, tx, ty);
    to_b16<NT, BN>(reinterpret_cast<uint64_t*>(s.pB), qb, tx, ty);
  } else {
    if (FORM == 0) to_a<NT, kT>(s.dA, d, tx, ty, n - c0, n - c0);
    to_b<NT, kT>(s.dB, d, tx, ty, n - c0, n - c0);
    to_a<NT, BM>(s.pA, qa, tx, ty, BM, kT);
    to_b<NT, BN>(s.pB, qb, tx, ty, kT, BN);
  }
  __syncthreads();
}

// The look-ahead block of round c (when c + 1 < N): the round's result on
// tile (e, e) = (c+1, c+1), min(X(e,e), P(e,c) (x) X(c,e)) (single_panel; the
// min3 form keeps the two panel products, P(e,c) and P(c,e), one per half),
// closed -- the next round's D* -- and stored.  The region that holds tile
// (e, e) does not store it this round.
template <int R, int NT>
__device__ __forceinline__ void lookahead_block(SmallSmem<R>& s, int32_t* D, int32_t n, int32_t c,
                                                bool small16_enabled, bool squaring) {
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4, th = (threadIdx.x & (NT / 2 - 1)) >> 4;
  const bool upper = threadIdx.x >= NT / 2;
  const int32_t c0 = c * kT, e0 = c0 + kT;
  Cells<NT, kT, kT> d, qa, qb, y;
  load_cells<NT, kT, kT>(D, n, c0, c0, tx, ty, d);
  load_cells<NT, kT, kT>(D, n, e0, c0, tx, ty, qa);
  load_cells<NT, kT, kT>(D, n, c0, e0, tx, ty, qb);
  load_cells<NT, kT, kT>(D, n, e0, e0, tx, ty, y);
  const uint32_t o = or_cells<NT, kT, kT>(d) | or_cells<NT, kT, kT>(qa) | or_cells<NT, kT, kT>(qb) |
                     or_cells<NT, kT, kT>(y);
  const uint32_t vote = form_vote<NT>(s.wflags, o);
  const bool pk = !(vote & 1u);
  const bool p16 = small16_enabled && !(vote & 2u);
  if (p16) {
    stage_round<R, NT, kT, kT, 2>(s, d, qa, qb, n, c0);
    single_panel<R, NT, kT, kT, 2>(s, qa, y, n, c0, e0);
    if (squaring && n - e0 >= kT)
      close_sq16<R, NT>(s, y);
    else
      close_regs<kT, NT>(y, s.lines, min(kT, n - e0));
  } else if (pk) {
    stage_round<R, NT, kT, kT, 1>(s, d, qa, qb, n, c0);
    single_panel<R, NT, kT, kT, 1>(s, qa, y, n, c0, e0);
    close_regs<kT, NT>(y, s.lines, min(kT, n - e0));
  } else {
    stage_round<R, NT, kT, kT, 0>(s, d, qa, qb, n, c0);
    Cells<NT / 2, kT, kT> p;
    if (!upper)
      from_a<NT / 2, kT>(s.pA, p, tx, th);  // X(e, c)
    else
      from_b<NT / 2, kT>(s.pB, p, tx, th);  // X(c, e)
    lookahead_products<R, NT, false>(s, p, y, upper, n, c0, e0);
    close_regs<kT, NT>(y, s.lines, min(kT, n - e0));
  }
  store_cells<NT, kT, kT>(D, n, e0, e0, tx, ty, y);
}

// Round c: block 0 is the look-ahead block (when c + 1 < N), block b > 0 the
// region b - 1 of M x M.  Tile (c, c) holds D* already (the previous
// round's look-ahead block stored it; k_small_close0 for round 0).  Every
// thread holds, in the NT-thread layout, its cells of tile (c,c), of the A
// panel X(I,c), of the B panel X(c,J) and of the region; they are loaded at
// once, then one block-wide vote picks the form.  pk16 and pk64 blocks run
// single_panel; min3 blocks (negative values) the two panel products of the
// three-phase order, one per half of the block (operands read back from
// the staging in the half's layout).
template <int R, int NT>
__global__ __launch_bounds__(NT) void k_small_round(int32_t* __restrict__ D, int32_t n, int32_t c, int32_t M,
                                                    int32_t N, int32_t flags) {
  __shared__ SmallSmem<R> s;
  const bool ahead = c + 1 < N;
  if (ahead && blockIdx.x == 0) {
    lookahead_block<R, NT>(s, D, n, c, (flags & 1) != 0, (flags & 2) != 0);
    return;
  }
  const int b = blockIdx.x - (ahead ? 1 : 0);
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4, th = (threadIdx.x & (NT / 2 - 1)) >> 4;
  const bool upper = threadIdx.x >= NT / 2;
  const int I = b / M, J = b - (b / M) * M;
  const int32_t c0 = c * kT, r0 = I * R, col0 = J * R;
  Cells<NT, kT, kT> d;
  Cells<NT, R, kT> qa;
  Cells<NT, kT, R> qb;
  Cells<NT, R, R> xc;
  load_cells<NT, kT, kT>(D, n, c0, c0, tx, ty, d);
  load_cells<NT, R, kT>(D, n, r0, c0, tx, ty, qa);
  load_cells<NT, kT, R>(D, n, c0, col0, tx, ty, qb);
  load_cells<NT, R, R>(D, n, r0, col0, tx, ty, xc);
  const uint32_t o = or_cells<NT, kT, kT>(d) | or_cells<NT, R, kT>(qa) | or_cells<NT, kT, R>(qb) |
                     or_cells<NT, R, R>(xc);
  const uint32_t vote = form_vote<NT>(s.wflags, o);
  const bool pk = !(vote & 1u);
  const bool p16 = (flags & 1) && !(vote & 2u);
  if (p16) {
    stage_round<R, NT, R, R, 2>(s, d, qa, qb, n, c0);
    single_panel<R, NT, R, R, 2>(s, qa, xc, n, c0, r0);
  } else if (pk) {
    stage_round<R, NT, R, R, 1>(s, d, qa, qb, n, c0);
    single_panel<R, NT, R, R, 1>(s, qa, xc, n, c0, r0);
  } else {
    stage_round<R, NT, R, R, 0>(s, d, qa, qb, n, c0);
    Cells<NT / 2, R, kT> pa;
    Cells<NT / 2, kT, R> pb;
    if (!upper)
      from_a<NT / 2, R>(s.pA, pa, tx, th);
    else
      from_b<NT / 2, R>(s.pB, pb, tx, th);
    small_products<R, NT, false>(s, pa, pb, xc, upper, n, c0, r0, col0);
  }
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

// FWX_SMALL_SQ=0 closes the look-ahead tile by the 32 Floyd-Warshall steps
// in every form (read at every call); else pk16 tiles close by squaring.
int small_sq() {
  const char* e = std::getenv("FWX_SMALL_SQ");
  return (e && e[0] == '0') ? 0 : 1;
}

}  // namespace

int enqueue_small_rounds(int32_t* d, int32_t n, hipStream_t st) {
  const int N = (n + kT - 1) / kT;
  const int R = small_region(n), NT = small_threads(), p16 = small16() | (small_sq() << 1);
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

int small_region_knob(int32_t n) {
  return small_region(n) | (small_threads() << 8) | (small16() << 16) | (small_sq() << 17);
}

}  // namespace fwx
