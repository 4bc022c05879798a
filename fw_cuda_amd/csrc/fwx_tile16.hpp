// fwx_tile16.hpp -- the packed 16-bit (pk16) min-plus tile product for gfx950.
//
//   C[BM x BN] <- min(C, A[BM x K] (x) B[K x BN])   on values clamped at CAP
//
// Encoding.  A value v >= 0 is carried as e(v) = min(v, CAP) in a 16-bit lane
// (CAP = 0x3BFF); A operands and C cells carry e(v) + BIAS (BIAS = 0x0400).
// Per k pair (k0, k1) and cell pair (j0, j1) of row i, ONE 64-bit add
//   (a_k0, a_k0, a_k1, a_k1) + (b_k0j0, b_k0j1, b_k1j0, b_k1j1)   (16-bit lanes)
// forms four sums (each below 0x7BFF: no lane carries), and ONE
// v_pk_minimum3_f16 folds them into the cell pair:
//   c = min3(c, sums(k0), sums(k1)) per lane.
// Every compared pattern lies in [BIAS, 0x7BFF], where binary16 numbers are
// positive, normal and finite and order exactly as their unsigned bit
// patterns, so the f16 minimum is the unsigned minimum (checked on the device
// by fw_cuda_amd/microbench/pk16_forms.hip over every pattern).  That is 2
// VALU ops per 4 relaxations, half of the pk64 form's 1 op per relaxation.
//
// Exactness.  Let h(x) = min(x, CAP).  For x, y >= 0, h(min(x, y)) =
// min(h(x), h(y)) and h(x + y) = h(h(x) + h(y)), so a product on encoded
// values yields h of the exact product: a result below CAP is the exact
// relaxation result; a result at CAP means the exact result is >= CAP, and
// the cell then keeps its old value (>= CAP as well, since its own encoding
// was CAP) -- store_c skips it.  So the matrix always holds walk lengths,
// and h(matrix) evolves exactly as h(FW state): after the run every cell
// below CAP holds its exact distance.  The caller checks for cells >= CAP
// afterwards and, if any, finishes with an exact int32 run on the matrix
// (fwx_blocked.hip).  Inputs must be >= 0 (the pk64 range check holds).
//
// Layout (256 threads, 16 x 16; thread (tx, ty) holds rows ty + 16 r and the
// 4-column groups 4 (tx + 16 g) .. +3, i.e. cell pairs 2 (tx + 16 g) + {0, 1}):
//   A chunk: row-major u32 e+BIAS duplicated into both lanes, row stride
//            tile_lda<KC> (the int32 product's A layout, same bytes);
//   B chunk: per k pair, the BN/2 column pairs in order, one 8-byte slot each
//            (low word: row k0's pair, high word: row k1's pair).
// A compute read is one 16-byte read per row (4 k, broadcast across the 16
// lanes of a row), a B read one 16-byte read per group and k pair (16 lanes
// on 256 consecutive bytes); the stash writes 16 bytes per lane on
// consecutive addresses -- conflict-free under gfx950's LDS bank model
// (tests/test_lds_banks.py).
#pragma once

#include "fwx_tile.hpp"

namespace fwx {

constexpr uint32_t kP16Bias = 0x0400u;
constexpr uint32_t kP16Cap = 0x3BFFu;
constexpr uint32_t kP16Bias2 = kP16Bias | (kP16Bias << 16);

// v_pk_minimum3_f16 (IEEE minimum of three, per 16-bit lane)
__device__ __forceinline__ uint32_t pkmin3_16(uint32_t a, uint32_t b, uint32_t c) {
  uint32_t r;
  asm("v_pk_minimum3_f16 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
  return r;
}

// two k steps on a cell pair: a = (a_k0, a_k0, a_k1, a_k1), b = (b_k0j0,
// b_k0j1, b_k1j0, b_k1j1), 16-bit lanes low to high
__device__ __forceinline__ uint32_t relax4_pk16(uint32_t c, uint64_t a, uint64_t b) {
  const uint64_t s = a + b;
  return pkmin3_16(c, (uint32_t)s, (uint32_t)(s >> 32));
}

// A operand: e(v) + BIAS in both lanes (v >= 0)
__device__ __forceinline__ uint32_t p16_dup(int32_t v) {
  return __umul24(min((uint32_t)v, kP16Cap), 0x10001u) + kP16Bias2;
}

// B operand pair: e(x) | e(y) << 16
__device__ __forceinline__ uint32_t p16_pair(int32_t x, int32_t y) {
  return min((uint32_t)x, kP16Cap) | (min((uint32_t)y, kP16Cap) << 16);
}

template <int BM, int BN, int KC>
struct alignas(16) TileSmem16 {
  alignas(16) uint32_t a[2][BM * tile_lda<KC>()];
  alignas(16) uint64_t b[2][(KC / 2) * (BN / 2)];
};

template <int BM, int BN, int KC, bool VEC, bool EDGE>
struct TileProduct16 : TileProduct<BM, BN, KC, VEC, 16, true, EDGE> {
  using Base = TileProduct<BM, BN, KC, VEC, 16, true, EDGE>;
  using Base::AG;
  using Base::APASS;
  using Base::AROWS;
  using Base::BPASS;
  using Base::BU;
  using Base::LDA;
  using Base::NT;
  using Base::ra;
  using Base::rb0;
  using Base::rb1;
  static constexpr int TM = BM / 16;  // rows per thread
  static constexpr int TG = BN / 64;  // 4-column groups per thread
  static constexpr int TP = 2 * TG;   // cell pairs per thread
  static_assert(BN % 64 == 0 && BM % 16 == 0, "pk16 block: 16 x 16 threads, 4-column groups");
  using Smem = TileSmem16<BM, BN, KC>;

  __device__ __forceinline__ void stash16(Smem& s, int buf, const View& A, const View& B) {
    const int tid = threadIdx.x;
    this->mask_edges(A, B);
#pragma unroll
    for (int p = 0; p < APASS; ++p) {
      const int r = p * AROWS + tid / AG;
      const int k = (tid % AG) * 4;
      if (BM % AROWS == 0 || r < BM) {
        const uint4 e = make_uint4(p16_dup(ra[p].x), p16_dup(ra[p].y), p16_dup(ra[p].z), p16_dup(ra[p].w));
        reinterpret_cast<uint4*>(s.a[buf])[(r * LDA + k) / 4] = e;
      }
    }
#pragma unroll
    for (int u = 0; u < BPASS; ++u) {
      const int idx = u * NT + tid;
      if (BU % NT == 0 || idx < BU) {
        const int pr = idx / (BN / 4);
        const int j = (idx % (BN / 4)) * 4;
        ulonglong2 v;
        v.x = (uint64_t)p16_pair(rb0[u].x, rb0[u].y) | ((uint64_t)p16_pair(rb1[u].x, rb1[u].y) << 32);
        v.y = (uint64_t)p16_pair(rb0[u].z, rb0[u].w) | ((uint64_t)p16_pair(rb1[u].z, rb1[u].w) << 32);
        reinterpret_cast<ulonglong2*>(s.b[buf])[(pr * (BN / 2) + j / 2) / 2] = v;
      }
    }
  }

  __device__ __forceinline__ static void compute16(const Smem& s, int buf, uint32_t (&acc)[TM][TP]) {
    const int tid = threadIdx.x;
    const int tx = tid % 16;
    const int ty = tid / 16;
    const uint32_t* As = s.a[buf];
    const uint64_t* Bs = s.b[buf];
#pragma unroll 1
    for (int q = 0; q < KC / 4; ++q) {
      uint64_t b[2][TP];
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int g = 0; g < TG; ++g) {
          const ulonglong2 v =
              *reinterpret_cast<const ulonglong2*>(&Bs[(2 * q + h) * (BN / 2) + 2 * (tx + 16 * g)]);
          b[h][2 * g] = v.x;
          b[h][2 * g + 1] = v.y;
        }
      // each A row is read one row ahead of its relaxations, and the
      // scheduling barriers keep that read in place: read just before its
      // use (the compiler's choice, one register set for every row), each
      // row's 4 TP instructions began with an LDS round trip.  Measured at
      // n=8192 on the packed rest kernel: 264.6 -> 255.7 us per launch.
      uint4 a = *reinterpret_cast<const uint4*>(&As[ty * LDA + 4 * q]);
#pragma unroll
      for (int r = 0; r < TM; ++r) {
        uint4 an;
        if (r + 1 < TM) an = *reinterpret_cast<const uint4*>(&As[(ty + 16 * (r + 1)) * LDA + 4 * q]);
        __builtin_amdgcn_sched_barrier(0);
        const uint64_t a01 = (uint64_t)a.x | ((uint64_t)a.y << 32);
        const uint64_t a23 = (uint64_t)a.z | ((uint64_t)a.w << 32);
#pragma unroll
        for (int p = 0; p < TP; ++p) {
          acc[r][p] = relax4_pk16(acc[r][p], a01, b[0][p]);
          acc[r][p] = relax4_pk16(acc[r][p], a23, b[1][p]);
        }
        __builtin_amdgcn_sched_barrier(0);
        if (r + 1 < TM) a = an;
      }
    }
  }

  // acc <- min(acc, A (x) B) over k in [0, depth), depth a multiple of KC.
  __device__ __forceinline__ void run16(Smem& s, const View& A, const View& B, int32_t depth,
                                        uint32_t (&acc)[TM][TP]) {
    const int nch = depth / KC;
    this->fetch(A, B, 0);
    stash16(s, 0, A, B);
    __syncthreads();
    for (int ch = 0; ch < nch; ++ch) {
      const int buf = ch & 1;
      if (ch + 1 < nch) this->fetch(A, B, (ch + 1) * KC);
      compute16(s, buf, acc);
      if (ch + 1 < nch) stash16(s, buf ^ 1, A, B);
      __syncthreads();
    }
  }

  // C cells: e(v) + BIAS per lane; out-of-range cells are INF (e = CAP).
  __device__ __forceinline__ static void load_c16(const View& C, uint32_t (&acc)[TM][TP]) {
    const int tid = threadIdx.x;
    const int tx = tid % 16;
    const int ty = tid / 16;
#pragma unroll
    for (int r = 0; r < TM; ++r)
#pragma unroll
      for (int g = 0; g < TG; ++g) {
        const int i = ty + 16 * r, j = 4 * (tx + 16 * g);
        int4 v;
        if (!EDGE && VEC) {
          v = *reinterpret_cast<const int4*>(C.at(i, j));
        } else if (!EDGE) {
          const int32_t* q = C.at(i, j);
          v = make_int4(q[0], q[1], q[2], q[3]);
        } else {
          v.x = C.in(i, j + 0) ? *C.at(i, j + 0) : kInf;
          v.y = C.in(i, j + 1) ? *C.at(i, j + 1) : kInf;
          v.z = C.in(i, j + 2) ? *C.at(i, j + 2) : kInf;
          v.w = C.in(i, j + 3) ? *C.at(i, j + 3) : kInf;
        }
        acc[r][2 * g] = p16_pair(v.x, v.y) + kP16Bias2;
        acc[r][2 * g + 1] = p16_pair(v.z, v.w) + kP16Bias2;
      }
  }

  // Stores every cell whose result is below CAP; a cell at CAP keeps its old
  // value (which is >= CAP, see the header).
  __device__ __forceinline__ static void store_c16(const View& C, const uint32_t (&acc)[TM][TP]) {
    uint32_t tid = threadIdx.x;
    asm("" : "+v"(tid));
    const int tx = tid % 16;
    const int ty = tid / 16;
#pragma unroll
    for (int r = 0; r < TM; ++r)
#pragma unroll
      for (int g = 0; g < TG; ++g) {
        const int i = ty + 16 * r, j = 4 * (tx + 16 * g);
        const uint32_t p0 = acc[r][2 * g] - kP16Bias2, p1 = acc[r][2 * g + 1] - kP16Bias2;
        const int32_t v0 = (int32_t)(p0 & 0xFFFFu), v1 = (int32_t)(p0 >> 16);
        const int32_t v2 = (int32_t)(p1 & 0xFFFFu), v3 = (int32_t)(p1 >> 16);
        const bool k0 = v0 < (int32_t)kP16Cap, k1 = v1 < (int32_t)kP16Cap;
        const bool k2 = v2 < (int32_t)kP16Cap, k3 = v3 < (int32_t)kP16Cap;
        if (!EDGE && VEC && k0 && k1 && k2 && k3) {
          *reinterpret_cast<int4*>(C.at(i, j)) = make_int4(v0, v1, v2, v3);
        } else {
          if (k0 && (!EDGE || C.in(i, j + 0))) *C.at(i, j + 0) = v0;
          if (k1 && (!EDGE || C.in(i, j + 1))) *C.at(i, j + 1) = v1;
          if (k2 && (!EDGE || C.in(i, j + 2))) *C.at(i, j + 2) = v2;
          if (k3 && (!EDGE || C.in(i, j + 3))) *C.at(i, j + 3) = v3;
        }
      }
  }

  // store_c16 for a tile that other workgroups update in the same launch:
  // every cell whose result is below CAP and differs from the value `was`
  // it was loaded with is lowered by an agent-scope atomic min, so no
  // workgroup's result overwrites a lower one.  Full tiles only.
  __device__ __forceinline__ static void min_c16(const View& C, const uint32_t (&acc)[TM][TP],
                                                 const uint32_t (&was)[TM][TP]) {
    static_assert(!EDGE, "min_c16: full tiles only");
    const int tid = threadIdx.x;
    const int tx = tid % 16;
    const int ty = tid / 16;
#pragma unroll
    for (int r = 0; r < TM; ++r)
#pragma unroll
      for (int p = 0; p < TP; ++p) {
        const int i = ty + 16 * r, j = 4 * (tx + 16 * (p / 2)) + 2 * (p % 2);
        const uint32_t v = acc[r][p] - kP16Bias2, w = acc[r][p] ^ was[r][p];
        const int32_t lo = (int32_t)(v & 0xFFFFu), hi = (int32_t)(v >> 16);
        if ((w & 0xFFFFu) != 0u && lo < (int32_t)kP16Cap)
          __hip_atomic_fetch_min(C.at(i, j), lo, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if ((w >> 16) != 0u && hi < (int32_t)kP16Cap)
          __hip_atomic_fetch_min(C.at(i, j + 1), hi, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
  }
};

// C <- min(C, A (x) B) in the pk16 form by one 256-thread workgroup; blocks
// wholly inside the matrix run the edge-free instantiation.
template <int BM, int BN, int KC, bool VEC>
__device__ __forceinline__ void tile_update16(TileSmem16<BM, BN, KC>& s, const View& A, const View& B,
                                              const View& C, int32_t depth) {
  using PF = TileProduct16<BM, BN, KC, VEC, false>;
  using PE = TileProduct16<BM, BN, KC, VEC, true>;
  uint32_t acc[PF::TM][PF::TP];
  if (C.rlim >= BM && C.clim >= BN && A.rlim >= BM && A.clim >= depth && B.rlim >= depth &&
      B.clim >= BN) {
    PF::load_c16(C, acc);
    PF p;
    p.run16(s, A, B, depth, acc);
    PF::store_c16(C, acc);
  } else {
    PE::load_c16(C, acc);
    PE p;
    p.run16(s, A, B, depth, acc);
    PE::store_c16(C, acc);
  }
}

// tile_update16 that also tells whether any cell of C changed (every thread
// of the block gets the same answer; one extra barrier).
template <int BM, int BN, int KC, bool VEC>
__device__ __forceinline__ bool tile_update16_changed(TileSmem16<BM, BN, KC>& s, const View& A, const View& B,
                                                      const View& C, int32_t depth) {
  using PF = TileProduct16<BM, BN, KC, VEC, false>;
  using PE = TileProduct16<BM, BN, KC, VEC, true>;
  uint32_t acc[PF::TM][PF::TP], was[PF::TM][PF::TP];
  const bool full = C.rlim >= BM && C.clim >= BN && A.rlim >= BM && A.clim >= depth && B.rlim >= depth &&
                    B.clim >= BN;
  if (full)
    PF::load_c16(C, acc);
  else
    PE::load_c16(C, acc);
  uint32_t diff = 0;
#pragma unroll
  for (int r = 0; r < PF::TM; ++r)
#pragma unroll
    for (int p = 0; p < PF::TP; ++p) was[r][p] = acc[r][p];
  if (full) {
    PF pf;
    pf.run16(s, A, B, depth, acc);
    PF::store_c16(C, acc);
  } else {
    PE pe;
    pe.run16(s, A, B, depth, acc);
    PE::store_c16(C, acc);
  }
#pragma unroll
  for (int r = 0; r < PF::TM; ++r)
#pragma unroll
    for (int p = 0; p < PF::TP; ++p) diff |= acc[r][p] ^ was[r][p];
  return __syncthreads_or(diff != 0) != 0;
}

// tile_update16_changed for one of several workgroups that update the same
// full C tile over disjoint parts of the depth in one launch: changed cells
// are lowered by atomic min (min_c16), so C ends at most min over the parts.
template <int BM, int BN, int KC, bool VEC>
__device__ __forceinline__ bool tile_update16_part(TileSmem16<BM, BN, KC>& s, const View& A, const View& B,
                                                   const View& C, int32_t depth) {
  using PF = TileProduct16<BM, BN, KC, VEC, false>;
  uint32_t acc[PF::TM][PF::TP], was[PF::TM][PF::TP];
  PF::load_c16(C, acc);
#pragma unroll
  for (int r = 0; r < PF::TM; ++r)
#pragma unroll
    for (int p = 0; p < PF::TP; ++p) was[r][p] = acc[r][p];
  PF pf;
  pf.run16(s, A, B, depth, acc);
  uint32_t diff = 0;
#pragma unroll
  for (int r = 0; r < PF::TM; ++r)
#pragma unroll
    for (int p = 0; p < PF::TP; ++p) diff |= acc[r][p] ^ was[r][p];
  PF::min_c16(C, acc, was);
  return __syncthreads_or(diff != 0) != 0;
}

// ---------------------------------------------------------------------------
// Pre-packed operands.  A stage's panels, packed once (k_pack16 in
// fwx_blocked.hip) in the exact byte layout of the LDS chunks above:
//   A: row i holds e(D[i][s0 + k]) + BIAS in both lanes for k < depth (u32,
//      row stride lda words) -- p16_dup of the int32 cell;
//   B: k pair p holds, per column pair (2c, 2c + 1), one 8-byte slot
//      p16_pair(row 2p) | p16_pair(row 2p + 1) << 32 (row stride ldb slots).
// The product then stages a chunk by plain 16-byte copies: no conversion
// instructions, and B takes half the bytes of its int32 cells.  Edge-free
// only: the caller guarantees whole tiles and a depth that is a multiple of
// KC.
struct PackedPanels {
  const uint32_t* a;  // (0, 0) of this tile's A rows
  int32_t lda;        // words
  const uint64_t* b;  // (0, 0) of this tile's B column pairs
  int32_t ldb;        // slots
};

template <int BM, int BN, int KC>
struct PackedProduct16 {
  using P = TileProduct16<BM, BN, KC, true, false>;
  using Smem = TileSmem16<BM, BN, KC>;
  static constexpr int NT = 256;
  static constexpr int LDA = tile_lda<KC>();
  static constexpr int AG = KC / 4;                     // 16-byte groups per A row
  static constexpr int AROWS = NT / AG;                 // A rows per pass
  static constexpr int APASS = BM / AROWS;
  static constexpr int BROW = BN / 4;                   // threads per k-pair row (2 slots each)
  static constexpr int BROWS = NT / BROW;               // k-pair rows per pass
  static constexpr int BPASS = (KC / 2) / BROWS;
  static_assert(APASS == 2 && BPASS == 1, "the 128 x 128, 16-deep shape");

  // acc <- min(acc, A (x) B) over k in [0, depth), depth a multiple of KC.
  // The staged groups are locals of this loop (two A groups and one B group
  // per thread); held as members of a product object, they were placed in
  // LDS by the compiler.
  __device__ __forceinline__ static void run(Smem& s, const PackedPanels& w, int32_t depth,
                                             uint32_t (&acc)[P::TM][P::TP]) {
    const int tid = threadIdx.x;
    const uint32_t oa = (uint32_t)((tid / AG) * w.lda + (tid % AG) * 4) * 4u;
    const uint32_t ob = (uint32_t)((tid / BROW) * w.ldb + (tid % BROW) * 2) * 8u;
    const int sa0 = ((tid / AG) * LDA + (tid % AG) * 4) / 4;
    const int sa1 = ((AROWS + tid / AG) * LDA + (tid % AG) * 4) / 4;
    const int sb = ((tid / BROW) * (BN / 2) + (tid % BROW) * 2) / 2;
    uint4 ra0, ra1, rb;
    auto fetch = [&](int32_t kc0) {
      ra0 = *at_bytes(reinterpret_cast<const uint4*>(w.a + kc0), oa);
      ra1 = *at_bytes(reinterpret_cast<const uint4*>(w.a + kc0 + AROWS * w.lda), oa);
      rb = *at_bytes(reinterpret_cast<const uint4*>(w.b + (kc0 / 2) * w.ldb), ob);
    };
    // the positions stash16 writes, the same bytes
    auto stash = [&](int buf) {
      reinterpret_cast<uint4*>(s.a[buf])[sa0] = ra0;
      reinterpret_cast<uint4*>(s.a[buf])[sa1] = ra1;
      reinterpret_cast<uint4*>(s.b[buf])[sb] = rb;
    };
    const int nch = depth / KC;
    fetch(0);
    stash(0);
    __syncthreads();
    for (int ch = 0; ch < nch; ++ch) {
      const int buf = ch & 1;
      if (ch + 1 < nch) fetch((ch + 1) * KC);
      P::compute16(s, buf, acc);
      if (ch + 1 < nch) stash(buf ^ 1);
      __syncthreads();
    }
  }
};

// C <- min(C, A (x) B) from packed panels by one 256-thread workgroup; C
// wholly inside the matrix.
template <int BM, int BN, int KC>
__device__ __forceinline__ void tile_update16_packed(TileSmem16<BM, BN, KC>& s, const PackedPanels& w,
                                                     const View& C, int32_t depth, int* __restrict__ capped) {
  using PP = PackedProduct16<BM, BN, KC>;
  uint32_t acc[PP::P::TM][PP::P::TP];
  PP::P::load_c16(C, acc);
  PP::run(s, w, depth, acc);
  PP::P::store_c16(C, acc);
  if (capped) {
    // a result lane at CAP is a cell left at its old value >= CAP
    constexpr uint32_t kTop = kP16Cap + kP16Bias;
    bool any = false;
#pragma unroll
    for (int r = 0; r < PP::P::TM; ++r)
#pragma unroll
      for (int p = 0; p < PP::P::TP; ++p) any |= (acc[r][p] & 0xFFFFu) == kTop || (acc[r][p] >> 16) == kTop;
    if (__ballot(any) != 0ull && (threadIdx.x & 63) == 0) atomicOr(capped, 1);
  }
}

// ---------------------------------------------------------------------------
// Closure of a T x T diagonal tile by min-plus squaring in the pk16 form.
//
// The tile sits in registers in close_regs's layout (thread (tx, ty) holds
// cells (ty + RS r, tx + 16 q), RS = NTH / 16).  When every cell lies in
// [0, kSq16Small) and the tile has no padding, y <- min(y, y (x) y) runs
// until a squaring changes no cell, at most log2(T) times: after s
// squarings a cell holds the shortest walk of at most 2^s edges over the
// tile's vertices; a squaring that changes nothing leaves y obeying every
// triangle inequality, i.e. y is the closure; after log2(T) squarings,
// walks of T edges cover every simple path.  With no negative cycle (all
// cells are >= 0) that is the closure the T Floyd-Warshall steps give, in
// 2 to log2(T) dependent tile products instead of T dependent steps.  Every
// value is a min over the tile's values and sums of two of them, so sums
// stay below 2 kSq16Small <= CAP and nothing saturates.  Otherwise the
// function returns false and the caller runs the T steps.
// A operand: (v + BIAS) in both lanes, row-major with stride T + 4 (16-byte
// rows); B operand: per k pair, column pairs (j, j + 16) with
// j = (cp % 16) + 32 (cp / 16) -- the pairs a thread's cells form -- one
// 8-byte slot each (low word row 2p, high word row 2p + 1).
constexpr uint32_t kSq16Small = 0x1000u;

template <int T, int NTH>
struct alignas(16) SqSmem16 {
  alignas(16) uint32_t a[T * (T + 4)];
  alignas(16) uint64_t b[(T / 2) * (T / 2)];
  uint32_t votes[2][NTH / 64];
};

// true if any thread of the block passes pred; one barrier.
template <int NTH>
__device__ __forceinline__ bool block_any(uint32_t* words, bool pred) {
  const bool w = __ballot(pred) != 0;
  if ((threadIdx.x & 63) == 0) words[threadIdx.x >> 6] = w ? 1u : 0u;
  __syncthreads();
  uint32_t any = 0;
#pragma unroll
  for (int i = 0; i < NTH / 64; ++i) any |= words[i];
  return any != 0;
}

template <int T, int NTH>
__device__ __forceinline__ bool close_sq16_regs(int32_t (&acc)[T / (NTH / 16)][T / 16], SqSmem16<T, NTH>& s,
                                                int kmax) {
  constexpr int RS = NTH / 16, TR = T / RS, TQ = T / 16, TP = T / 32, LDA = T + 4;
  constexpr int LOG2T = T == 32 ? 5 : T == 64 ? 6 : 7;
  static_assert(T == 32 || T == 64 || T == 128, "tile width");
  static_assert(RS % 16 == 0 && T % RS == 0, "layout must tile T");
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  uint32_t o = 0;
#pragma unroll
  for (int r = 0; r < TR; ++r)
#pragma unroll
    for (int q = 0; q < TQ; ++q) o |= (uint32_t)acc[r][q];
  if (kmax < T || block_any<NTH>(s.votes[0], o >= kSq16Small)) return false;
  uint32_t* bw = reinterpret_cast<uint32_t*>(s.b);
  for (int it = 0; it < LOG2T; ++it) {
    // stage y as both operands (the previous squaring's reads ended before
    // its vote barrier)
#pragma unroll
    for (int r = 0; r < TR; ++r) {
      const int row = ty + RS * r;
#pragma unroll
      for (int q = 0; q < TQ; ++q) s.a[row * LDA + tx + 16 * q] = ((uint32_t)acc[r][q] + kP16Bias) * 0x10001u;
#pragma unroll
      for (int qp = 0; qp < TP; ++qp)
        bw[((row >> 1) * (T / 2) + tx + 16 * qp) * 2 + (row & 1)] =
            (uint32_t)acc[r][2 * qp] | ((uint32_t)acc[r][2 * qp + 1] << 16);
    }
    __syncthreads();
    uint32_t c[TR][TP], was[TR][TP];
#pragma unroll
    for (int r = 0; r < TR; ++r)
#pragma unroll
      for (int qp = 0; qp < TP; ++qp) {
        c[r][qp] = ((uint32_t)acc[r][2 * qp] | ((uint32_t)acc[r][2 * qp + 1] << 16)) + kP16Bias2;
        was[r][qp] = c[r][qp];
      }
    // c <- min(c, y (x) y): groups of 4 k, the next group read while the
    // current one is computed
    uint4 a[2][TR];
    uint64_t b[2][2][TP];
    auto load = [&](int buf, int g) {
#pragma unroll
      for (int r = 0; r < TR; ++r) a[buf][r] = *reinterpret_cast<const uint4*>(&s.a[(ty + RS * r) * LDA + 4 * g]);
#pragma unroll
      for (int qp = 0; qp < TP; ++qp) {
        b[buf][0][qp] = s.b[(2 * g) * (T / 2) + tx + 16 * qp];
        b[buf][1][qp] = s.b[(2 * g + 1) * (T / 2) + tx + 16 * qp];
      }
    };
    load(0, 0);
#pragma unroll 2
    for (int g = 0; g < T / 4; ++g) {
      const int cur = g & 1;
      if (g + 1 < T / 4) load(cur ^ 1, g + 1);
#pragma unroll
      for (int r = 0; r < TR; ++r) {
        const uint64_t a01 = (uint64_t)a[cur][r].x | ((uint64_t)a[cur][r].y << 32);
        const uint64_t a23 = (uint64_t)a[cur][r].z | ((uint64_t)a[cur][r].w << 32);
#pragma unroll
        for (int qp = 0; qp < TP; ++qp) {
          c[r][qp] = relax4_pk16(c[r][qp], a01, b[cur][0][qp]);
          c[r][qp] = relax4_pk16(c[r][qp], a23, b[cur][1][qp]);
        }
      }
    }
    uint32_t changed = 0;
#pragma unroll
    for (int r = 0; r < TR; ++r)
#pragma unroll
      for (int qp = 0; qp < TP; ++qp) {
        changed |= c[r][qp] ^ was[r][qp];
        const uint32_t v = c[r][qp] - kP16Bias2;
        acc[r][2 * qp] = (int32_t)(v & 0xFFFFu);
        acc[r][2 * qp + 1] = (int32_t)(v >> 16);
      }
    if (!block_any<NTH>(s.votes[(it + 1) & 1], changed != 0)) break;
  }
  return true;
}

// Close the T x T tile at V in place: by squaring where close_sq16_regs
// applies, else by the T Floyd-Warshall steps (which borrow the A staging
// buffer for their lines).  Cells outside the matrix read as INF and are
// not stored.
template <int T, int NTH>
__device__ __forceinline__ void close_tile_sq(const View V, SqSmem16<T, NTH>& s) {
  constexpr int RS = NTH / 16, TR = T / RS, TQ = T / 16;
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  int32_t acc[TR][TQ];
#pragma unroll
  for (int r = 0; r < TR; ++r)
#pragma unroll
    for (int q = 0; q < TQ; ++q)
      acc[r][q] = V.in(ty + RS * r, tx + 16 * q) ? *V.at(ty + RS * r, tx + 16 * q) : kInf;
  const int kmax = min(T, V.rlim);
  if (!close_sq16_regs<T, NTH>(acc, s, kmax)) close_regs<T, NTH>(acc, reinterpret_cast<int32_t*>(s.a), kmax);
#pragma unroll
  for (int r = 0; r < TR; ++r)
#pragma unroll
    for (int q = 0; q < TQ; ++q)
      if (V.in(ty + RS * r, tx + 16 * q)) *V.at(ty + RS * r, tx + 16 * q) = acc[r][q];
}

}  // namespace fwx
