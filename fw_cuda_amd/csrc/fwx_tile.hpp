// fwx_tile.hpp -- register-blocked min-plus tile product for gfx950.
//
//   C[BM x BN] <- min(C, A[BM x K] (x) B[K x BN])      ((x) = min-plus product)
//
// One workgroup of NTX x 16 threads (256 or 512: one or two wave64s per
// SIMD) owns the C block in VGPRs: thread (tx, ty) = (tid % NTX, tid / NTX)
// holds cells in rows ty + 16 r and in groups of up to four adjacent columns
// (colmap below).  LDS accesses are bank-conflict-free under gfx950's bank
// model as measured per access pattern (fw_cuda_amd/microbench/lds_banks.hip;
// tests/test_lds_banks.py walks every access of every shape): a wave's tx
// lanes read consecutive 64-bit B pairs, its ty rows read A as 16-byte
// broadcasts, and the stash writes whole 16-byte A groups and 8-byte B pairs.
//
// K is walked in chunks of KC, double-buffered through LDS with one barrier
// per chunk.  A is staged row-major (k contiguous), B pair-interleaved
// (B[k][j], B[k+1][j] adjacent), so both operands of a k pair are one 64-bit
// register each -- the layout the pk64 form needs (fwx_device.hpp).  The
// arithmetic form (PK64 or min3) is a template parameter chosen per call by
// the host from a range check of the input (fwx_range.hip): carrying both
// forms in one kernel costs ~60 VGPRs of the min3 form's temporaries.
//
// Out-of-range operand cells (row/col/k >= n) are staged as INF and C cells
// outside the matrix are never stored, so ragged n needs no padding buffer.
#pragma once

#include <type_traits>

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

// base + an unsigned 32-bit byte offset.  With a uniform base this is the
// shape the global load/store "saddr" form takes (64-bit base in SGPRs, one
// 32-bit VGPR offset), instead of a 64-bit VGPR address per access.  Every
// per-thread offset used with it is below 2^32 bytes (a window row of at most
// 32769 int32 cells times at most 128 rows).
template <class T>
__device__ __forceinline__ T* at_bytes(T* base, uint32_t bytes) {
  using C = typename std::conditional<std::is_const<T>::value, const char, char>::type;
  return reinterpret_cast<T*>(reinterpret_cast<C*>(base) + bytes);
}

// The thread's index within its block.  OPAQUE reads it through an empty
// volatile asm at each use site: values derived from it are then recomputed
// where they are needed instead of being treated as invariant across an
// enclosing loop.  The persistent batch kernel runs many products in a loop;
// with the plain index the compiler hoisted every per-thread offset of the
// product out of that loop, kept them live across it and spilled VGPRs
// inside the product's inner loops.
template <bool OPAQUE>
__device__ __forceinline__ int block_tid() {
  uint32_t t = threadIdx.x;
  if (OPAQUE) asm volatile("" : "+v"(t));
  return (int)t;
}

// Row stride (dwords) of the staged A chunk.  gfx950's LDS has 32 banks of
// 4 bytes; a 16-byte write serves 8 lanes per cycle over 128 bytes, which the
// A stash fills with one 32-deep row (8 groups of 4 k) or two 16-deep rows:
// 36 dwords keep every 32-deep row start 16-byte aligned, and 16-deep rows
// stay unpadded so that two consecutive rows are 128 contiguous bytes.
template <int KC>
constexpr int tile_lda() {
  return KC == 16 ? 16 : KC + 4;
}

// LDS staging of one product: two buffers of an A chunk (row-major, row
// stride tile_lda) and a B chunk (pair-interleaved).  Both are 16-byte
// aligned so that the stash stores 16- and 8-byte words (4-byte pairs on
// 8-byte slots conflict 2-way).
template <int BM, int BN, int KC>
struct alignas(16) TileSmem {
  alignas(16) int32_t a[2][BM * tile_lda<KC>()];
  alignas(16) int32_t b[2][(KC / 2) * BN * 2];
};

// VEC: 16-byte global loads (needs ld % 4 == 0, 16-byte aligned base and
// c0 % 4 == 0, so a 4-wide group is wholly in or wholly out of range).
// NTX threads across columns x 16 across rows: 256 (NTX=16) or 512 (NTX=32).
// EDGE=false: the caller guarantees that the C block and every operand cell
// of the whole depth lie inside the matrix (tile_update checks it), so the
// clamping and masking code is compiled out.
// OPQ: the thread index is read with block_tid<true> (for products run in a
// loop of a persistent kernel).
template <int BM, int BN, int KC, bool VEC, int NTX, bool PK64, bool EDGE = true, bool OPQ = false>
struct TileProduct {
  static constexpr int NT = NTX * 16;
  static constexpr int NW = NT / 64;  // waves
  static constexpr int TM = BM / 16;
  static constexpr int TN = BN / NTX;
  static constexpr int LDA = tile_lda<KC>();
  // Column placement.  Thread tx owns groups of G = min(4, TN) adjacent
  // columns: cell slot c is column colmap(tx, c) = G*(tx + NTX*(c/G)) + c%G.
  // In LDS, B's pair row stores column j at pair slot bpos(j) =
  // (j%G)*(BN/G) + j/G, so a compute read (fixed c, all tx) and a stash write
  // (fixed column offset e, consecutive 4-column groups) both touch
  // consecutive 8-byte slots across lanes -- conflict-free either way.
  static constexpr int G = TN >= 4 ? 4 : TN;
  __device__ __forceinline__ static int colmap(int tx, int c) { return G * (tx + NTX * (c / G)) + c % G; }
  __device__ __forceinline__ static int bpos(int j) { return (j % G) * (BN / G) + j / G; }
  // 8-byte slot of pair (pr, j) in the B buffer.  An 8-byte write serves 16
  // lanes per cycle over 128 bytes; when a pair row holds fewer than 16
  // staging units (BN = 32), the 16 lanes span two pair rows 256 bytes apart,
  // which would hit the same banks, so odd pair rows are rotated by one slot.
  static constexpr int SKEW = (BN / 4 < 16) ? 1 : 0;
  __device__ __forceinline__ static int bslot(int pr, int j) {
    return pr * BN + (SKEW ? ((bpos(j) + (pr & 1)) & (BN - 1)) : bpos(j));
  }
  static_assert(BM % 16 == 0 && BN % NTX == 0, "block must divide over the threads");
  static_assert(KC % 4 == 0, "chunk must hold whole k quads");
  // A staging: KC/4 int4 groups per row.
  static constexpr int AG = KC / 4;
  static constexpr int AROWS = NT / AG;  // rows per pass
  static constexpr int APASS = (BM + AROWS - 1) / AROWS;
  // B staging: units of (k pair, 4 columns).
  static constexpr int BU = (KC / 2) * (BN / 4);
  static constexpr int BPASS = (BU + NT - 1) / NT;

  using Smem = TileSmem<BM, BN, KC>;

  int4 ra[APASS];
  int4 rb0[BPASS], rb1[BPASS];
  bool full_;
  int32_t kc_;

  // global -> registers for the chunk starting at k offset kc0.  The loads
  // are unconditional: out-of-range rows/columns are clamped into the matrix
  // and replaced by INF in stash(), so no branch merges around the loads and
  // they stay in flight while the previous chunk is computed.  When the chunk
  // is wholly inside the matrix (the common case), every address is a
  // uniform row base (SGPRs) plus one per-thread offset plus an immediate.
  __device__ __forceinline__ void fetch(const View& A, const View& B, int32_t kc0) {
    const int tid = block_tid<OPQ>();
    static_assert(NT % (BN / 4) == 0, "B staging rows must be uniform per pass");
    constexpr int BROWS = NT / (BN / 4);  // k pairs per B pass
    full_ = !EDGE || (A.rlim >= BM && A.clim >= kc0 + KC && B.rlim >= kc0 + KC && B.clim >= BN);
    kc_ = kc0;
    if (VEC) {
      const int32_t* qa[APASS];
      const int32_t* qb[BPASS];
      if (full_) {
        const uint32_t oa = (uint32_t)((tid / AG) * A.ld + (tid % AG) * 4) * 4u;
#pragma unroll
        for (int p = 0; p < APASS; ++p)
          if (BM % AROWS == 0 || p * AROWS + tid / AG < BM)
            ra[p] = *at_bytes(reinterpret_cast<const int4*>(A.p + kc0 + (p * AROWS) * A.ld), oa);
        const uint32_t ob = (uint32_t)((tid / (BN / 4)) * 2 * B.ld + (tid % (BN / 4)) * 4) * 4u;
#pragma unroll
        for (int u = 0; u < BPASS; ++u) {
          if (BU % NT == 0 || u * NT + tid < BU) {
            const int32_t* pb = B.p + (kc0 + u * BROWS * 2) * B.ld;
            rb0[u] = *at_bytes(reinterpret_cast<const int4*>(pb), ob);
            rb1[u] = *at_bytes(reinterpret_cast<const int4*>(pb + B.ld), ob);
          }
        }
        return;
      } else {
#pragma unroll
        for (int p = 0; p < APASS; ++p) {
          const int r = min(p * AROWS + tid / AG, A.rlim - 1);
          const int k = min(kc0 + (tid % AG) * 4, A.clim - 4);
          qa[p] = A.at(r, k);
        }
#pragma unroll
        for (int u = 0; u < BPASS; ++u) {
          const int idx = u * NT + tid;
          const int k = min(kc0 + 2 * (idx / (BN / 4)), B.rlim - 1);
          const int j = min((idx % (BN / 4)) * 4, B.clim - 4);
          qb[u] = B.at(k, j);
        }
      }
#pragma unroll
      for (int p = 0; p < APASS; ++p)
        if (BM % AROWS == 0 || p * AROWS + tid / AG < BM) ra[p] = *reinterpret_cast<const int4*>(qa[p]);
#pragma unroll
      for (int u = 0; u < BPASS; ++u) {
        if (BU % NT == 0 || u * NT + tid < BU) {
          rb0[u] = *reinterpret_cast<const int4*>(qb[u]);
          const int idx = u * NT + tid;
          const bool second = full_ || kc0 + 2 * (idx / (BN / 4)) + 1 < B.rlim;
          rb1[u] = *reinterpret_cast<const int4*>(qb[u] + (second ? B.ld : 0));
        }
      }
    } else {
#pragma unroll
      for (int p = 0; p < APASS; ++p) {
        const int r = min(p * AROWS + tid / AG, A.rlim - 1);
        const int k = kc0 + (tid % AG) * 4;
        const int32_t* q = A.p + r * A.ld;
        if (BM % AROWS == 0 || p * AROWS + tid / AG < BM) {
          ra[p].x = q[min(k + 0, A.clim - 1)];
          ra[p].y = q[min(k + 1, A.clim - 1)];
          ra[p].z = q[min(k + 2, A.clim - 1)];
          ra[p].w = q[min(k + 3, A.clim - 1)];
        }
      }
#pragma unroll
      for (int u = 0; u < BPASS; ++u) {
        const int idx = u * NT + tid;
        if (BU % NT == 0 || idx < BU) {
          const int k0 = min(kc0 + 2 * (idx / (BN / 4)), B.rlim - 1);
          const int k1 = min(kc0 + 2 * (idx / (BN / 4)) + 1, B.rlim - 1);
          const int j = (idx % (BN / 4)) * 4;
          const int32_t* q0 = B.p + k0 * B.ld;
          const int32_t* q1 = B.p + k1 * B.ld;
          rb0[u].x = q0[min(j + 0, B.clim - 1)];
          rb0[u].y = q0[min(j + 1, B.clim - 1)];
          rb0[u].z = q0[min(j + 2, B.clim - 1)];
          rb0[u].w = q0[min(j + 3, B.clim - 1)];
          rb1[u].x = q1[min(j + 0, B.clim - 1)];
          rb1[u].y = q1[min(j + 1, B.clim - 1)];
          rb1[u].z = q1[min(j + 2, B.clim - 1)];
          rb1[u].w = q1[min(j + 3, B.clim - 1)];
        }
      }
    }
  }

  __device__ __forceinline__ static int32_t keep(bool ok, int32_t v) { return ok ? v : kInf; }

  // Cells that fetch() clamped into the matrix become INF.
  __device__ __forceinline__ void mask_edges(const View& A, const View& B) {
    const int tid = block_tid<OPQ>();
    if (!full_) {
#pragma unroll
      for (int p = 0; p < APASS; ++p) {
        const int r = p * AROWS + tid / AG;
        const int k = kc_ + (tid % AG) * 4;
        const bool rok = r < A.rlim;
        ra[p].x = keep(rok && k + 0 < A.clim, ra[p].x);
        ra[p].y = keep(rok && k + 1 < A.clim, ra[p].y);
        ra[p].z = keep(rok && k + 2 < A.clim, ra[p].z);
        ra[p].w = keep(rok && k + 3 < A.clim, ra[p].w);
      }
#pragma unroll
      for (int u = 0; u < BPASS; ++u) {
        const int idx = u * NT + tid;
        const int k0 = kc_ + 2 * (idx / (BN / 4));
        const int j = (idx % (BN / 4)) * 4;
        const bool r0 = k0 < B.rlim, r1 = k0 + 1 < B.rlim;
        rb0[u].x = keep(r0 && j + 0 < B.clim, rb0[u].x);
        rb0[u].y = keep(r0 && j + 1 < B.clim, rb0[u].y);
        rb0[u].z = keep(r0 && j + 2 < B.clim, rb0[u].z);
        rb0[u].w = keep(r0 && j + 3 < B.clim, rb0[u].w);
        rb1[u].x = keep(r1 && j + 0 < B.clim, rb1[u].x);
        rb1[u].y = keep(r1 && j + 1 < B.clim, rb1[u].y);
        rb1[u].z = keep(r1 && j + 2 < B.clim, rb1[u].z);
        rb1[u].w = keep(r1 && j + 3 < B.clim, rb1[u].w);
      }
    }
  }

  // registers -> LDS buffer `buf`.  Cells that fetch() clamped become INF.
  __device__ __forceinline__ void stash(Smem& s, int buf, const View& A, const View& B) {
    const int tid = block_tid<OPQ>();
    mask_edges(A, B);
#pragma unroll
    for (int p = 0; p < APASS; ++p) {
      const int r = p * AROWS + tid / AG;
      const int k = (tid % AG) * 4;
      if (BM % AROWS == 0 || r < BM)
        reinterpret_cast<int4*>(s.a[buf])[(r * LDA + k) / 4] = ra[p];
    }
#pragma unroll
    for (int u = 0; u < BPASS; ++u) {
      const int idx = u * NT + tid;
      if (BU % NT == 0 || idx < BU) {
        const int pr = idx / (BN / 4);
        const int j = (idx % (BN / 4)) * 4;
        int2* slots = reinterpret_cast<int2*>(s.b[buf]);
        slots[bslot(pr, j + 0)] = make_int2(rb0[u].x, rb1[u].x);
        slots[bslot(pr, j + 1)] = make_int2(rb0[u].y, rb1[u].y);
        slots[bslot(pr, j + 2)] = make_int2(rb0[u].z, rb1[u].z);
        slots[bslot(pr, j + 3)] = make_int2(rb0[u].w, rb1[u].w);
      }
    }
  }

  __device__ __forceinline__ static void compute(const Smem& s, int buf, int32_t (&acc)[TM][TN]) {
    compute(s, buf, acc, block_tid<OPQ>());
  }

  __device__ __forceinline__ static void compute(const Smem& s, int buf, int32_t (&acc)[TM][TN], int tid) {
    const int tx = tid % NTX;
    const int ty = tid / NTX;
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
        b0[c] = Bs[bslot(2 * q, colmap(tx, c))];
        b1[c] = Bs[bslot(2 * q + 1, colmap(tx, c))];
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
    stash(s, 0, A, B);
    __syncthreads();
    for (int ch = 0; ch < nch; ++ch) {
      const int buf = ch & 1;
      if (ch + 1 < nch) fetch(A, B, (ch + 1) * KC);
      compute(s, buf, acc, block_tid<OPQ>());
      if (ch + 1 < nch) stash(s, buf ^ 1, A, B);
      __syncthreads();
    }
  }

  __device__ __forceinline__ static bool full(const View& C) {
    return !EDGE || (C.rlim >= BM && C.clim >= BN);
  }

  __device__ __forceinline__ static void load_c(const View& C, int32_t (&acc)[TM][TN], int tid) {
    const int tx = tid % NTX;
    const int ty = tid / NTX;
    if (full(C)) {
      const uint32_t oc = (uint32_t)(ty * C.ld + G * tx) * 4u;
#pragma unroll
      for (int r = 0; r < TM; ++r) {
        const int32_t* pr = C.p + (16 * r) * C.ld;
#pragma unroll
        for (int c = 0; c < TN; ++c) acc[r][c] = *at_bytes(pr + colmap(0, c), oc);
      }
    } else {
#pragma unroll
      for (int r = 0; r < TM; ++r)
#pragma unroll
        for (int c = 0; c < TN; ++c)
          acc[r][c] = C.in(ty + 16 * r, colmap(tx, c)) ? *C.at(ty + 16 * r, colmap(tx, c)) : kInf;
    }
  }

  // The thread index is passed through an empty asm so that the store
  // addresses are recomputed here rather than shared with load_c: shared,
  // they would hold one 64-bit VGPR address per row across the whole depth
  // loop (16 VGPRs at 8 rows).
  __device__ __forceinline__ static void store_c(const View& C, const int32_t (&acc)[TM][TN]) {
    uint32_t tid = threadIdx.x;
    asm("" : "+v"(tid));
    const int tx = tid % NTX;
    const int ty = tid / NTX;
    if (full(C)) {
      const uint32_t oc = (uint32_t)(ty * C.ld + G * tx) * 4u;
#pragma unroll
      for (int r = 0; r < TM; ++r) {
        int32_t* pr = C.p + (16 * r) * C.ld;
#pragma unroll
        for (int c = 0; c < TN; ++c) *at_bytes(pr + colmap(0, c), oc) = acc[r][c];
      }
    } else {
#pragma unroll
      for (int r = 0; r < TM; ++r)
#pragma unroll
        for (int c = 0; c < TN; ++c)
          if (C.in(ty + 16 * r, colmap(tx, c))) *C.at(ty + 16 * r, colmap(tx, c)) = acc[r][c];
    }
  }
};

// C <- min(C, A (x) B) over k in [0, depth) by one workgroup.  Blocks whose
// C block and operands lie wholly inside the matrix (every block of an
// n % BM == 0 matrix) run the edge-free instantiation, the others the
// general one; the branch is uniform across the block.
template <int BM, int BN, int KC, bool VEC, int NTX, bool PK64, bool OPQ = false>
__device__ __forceinline__ void tile_update(TileSmem<BM, BN, KC>& s, const View& A, const View& B,
                                            const View& C, int32_t depth) {
  using PF = TileProduct<BM, BN, KC, VEC, NTX, PK64, false, OPQ>;
  using PE = TileProduct<BM, BN, KC, VEC, NTX, PK64, true, OPQ>;
  int32_t acc[PF::TM][PF::TN];
  if (C.rlim >= BM && C.clim >= BN && A.rlim >= BM && A.clim >= depth && B.rlim >= depth &&
      B.clim >= BN) {
    PF::load_c(C, acc, block_tid<OPQ>());
    PF p;
    p.run(s, A, B, depth, acc);
    PF::store_c(C, acc);
  } else {
    PE::load_c(C, acc, block_tid<OPQ>());
    PE p;
    p.run(s, A, B, depth, acc);
    PE::store_c(C, acc);
  }
}

// Close a T x T diagonal tile in place (Floyd-Warshall inside the tile), one
// workgroup of NTH threads (16 columns x RS = NTH/16 rows of threads).
// Thread (tx, ty) keeps cells (ty + RS r, tx + 16 q) in VGPRs.  Two k steps
// per barrier: the owners of rows k, k+1 and columns k, k+1 publish them to
// a double-buffered LDS line; every thread then forms
//   col'[i] = min(D[i][k+1], D[i][k] + D[k][k+1])   (column k+1 after step k)
//   row'[j] = min(D[k+1][j], D[k+1][k] + D[k][j])   (row k+1 after step k)
//   D[i][j] = min3(D[i][j], D[i][k] + D[k][j], col'[i] + row'[j])
// which is exactly steps k and k+1 of the sequential loop.  Line buffer p&1
// is rewritten at pair p+2, after every thread passed barrier p+1, so one
// barrier per pair suffices.  Steps with k >= n (padding) are skipped, so
// padded vertices are never intermediates.  More threads mean fewer cells
// per thread and more waves per SIMD to hide the barrier and LDS latency of
// the T dependent steps.
// `lines` is LDS scratch of 8 T ints (the caller's, so a kernel that also
// runs tile products can lend their staging buffer).
// close_regs runs the steps on a tile already in the threads' registers
// (acc[r][q] = cell (ty + RS r, tx + 16 q)), for k < kmax.
template <int T, int NTH = 256>
__device__ __forceinline__ void close_regs(int32_t (&acc)[T / (NTH / 16)][T / 16], int32_t* lines, int kmax) {
  constexpr int RS = NTH / 16;  // thread rows
  constexpr int TR = T / RS;    // register rows per thread
  constexpr int TQ = T / 16;    // register columns per thread
  static_assert(NTH % 32 == 0 && T % RS == 0 && RS % 16 == 0, "layout must tile T");
  int32_t(*rows)[2][T] = reinterpret_cast<int32_t(*)[2][T]>(lines);          // [buf][k, k+1][j]
  int32_t(*cols)[2][T] = reinterpret_cast<int32_t(*)[2][T]>(lines + 4 * T);  // [buf][k, k+1][i]
  const int tx = threadIdx.x & 15;
  const int ty = threadIdx.x >> 4;
#pragma unroll
  for (int qk = 0; qk < TQ; ++qk) {
    // steps k = 16 qk + k16: column k is register column qk of threads
    // tx = k16, row k is register row rk of threads ty = k % RS (RS is a
    // multiple of 16, so rk is the same for the 16 steps)
    constexpr int RPQ = RS / 16;  // register-column blocks per register row
    const int rk = qk / RPQ;
    for (int k16 = 0; k16 < 16; k16 += 2) {
      const int k = qk * 16 + k16;
      if (k >= kmax) break;
      const int b = (k >> 1) & 1;
      const int tyk = k % RS;
      if (ty == tyk || ty == tyk + 1) {
        const int h = ty - tyk;
#pragma unroll
        for (int q = 0; q < TQ; ++q) rows[b][h][tx + 16 * q] = acc[rk][q];
      }
      if (tx == k16 || tx == k16 + 1) {
        const int h = tx - k16;
#pragma unroll
        for (int r = 0; r < TR; ++r) cols[b][h][ty + RS * r] = acc[r][qk];
      }
      __syncthreads();
      if (k + 1 < kmax) {
        const int32_t dkk1 = rows[b][0][k + 1];  // D[k][k+1]
        const int32_t dk1k = rows[b][1][k];      // D[k+1][k]
        int32_t rw[TQ], rw1[TQ], cl[TR], cl1[TR];
#pragma unroll
        for (int q = 0; q < TQ; ++q) {
          rw[q] = rows[b][0][tx + 16 * q];
          rw1[q] = min(rows[b][1][tx + 16 * q], wadd(dk1k, rw[q]));
        }
#pragma unroll
        for (int r = 0; r < TR; ++r) {
          cl[r] = cols[b][0][ty + RS * r];
          cl1[r] = min(cols[b][1][ty + RS * r], wadd(cl[r], dkk1));
        }
#pragma unroll
        for (int r = 0; r < TR; ++r)
#pragma unroll
          for (int q = 0; q < TQ; ++q) acc[r][q] = vmin3(acc[r][q], wadd(cl[r], rw[q]), wadd(cl1[r], rw1[q]));
      } else {
#pragma unroll
        for (int r = 0; r < TR; ++r) {
          const int32_t clr = cols[b][0][ty + RS * r];
#pragma unroll
          for (int q = 0; q < TQ; ++q) acc[r][q] = min(acc[r][q], wadd(clr, rows[b][0][tx + 16 * q]));
        }
      }
    }
  }
}

template <int T, int NTH = 256>
__device__ __forceinline__ void close_tile_in(const View V, int32_t* lines) {
  constexpr int RS = NTH / 16;
  constexpr int TR = T / RS;
  constexpr int TQ = T / 16;
  const int tx = threadIdx.x & 15;
  const int ty = threadIdx.x >> 4;
  int32_t acc[TR][TQ];
#pragma unroll
  for (int r = 0; r < TR; ++r)
#pragma unroll
    for (int q = 0; q < TQ; ++q)
      acc[r][q] = V.in(ty + RS * r, tx + 16 * q) ? *V.at(ty + RS * r, tx + 16 * q) : kInf;
  close_regs<T, NTH>(acc, lines, min(T, V.rlim));
#pragma unroll
  for (int r = 0; r < TR; ++r)
#pragma unroll
    for (int q = 0; q < TQ; ++q)
      if (V.in(ty + RS * r, tx + 16 * q)) *V.at(ty + RS * r, tx + 16 * q) = acc[r][q];
}

template <int T, int NTH = 256>
__device__ __forceinline__ void close_tile(const View V) {
  __shared__ int32_t lines[8 * T];
  close_tile_in<T, NTH>(V, lines);
}

}  // namespace fwx
