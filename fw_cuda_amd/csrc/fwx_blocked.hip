// fwx_blocked.hip -- blocked and multi-layer Floyd-Warshall for gfx950.
//
// Reference behaviour replaced: fwTilingGPU (src/fw_tiling_gpu.cu:253-284)
// and fwMultiLayerTilingGPU (src/fw_multi_layer_tiling_gpu.cu:507-577).  Both
// compute the all-pairs shortest paths of the INF-completed graph in place;
// with no negative cycle and no int32 overflow that matrix is unique, so the
// blocked orders below give bit-for-bit the result of the k-i-j loop
// (src/fw_baseline_cpu.cpp:3-16).
//
// The matrix is cut into T x T k-tiles (T = 32, 64, 128; N = ceil(n/T)).
// A stage is kappa consecutive k-tiles K = {kb .. kb+B-1}.  Per stage:
//
//   for each k-tile c in K, in order (the "lead" rounds):
//     diag  : close tile (c,c) with T dependent k steps      1 workgroup
//     panel : row c and column c tiles  C <- min(C, Dcc+ (x) C) or
//             min(C, C (x) Dcc+)  (Dcc+ is closed, so one product is exact)
//     cross : every other tile with a row or column in K
//             C <- min(C, D[i,c] (x) D[c,j])
//   rest  : every tile outside K's rows and columns
//             C <- min(C, D[I,K] (x) D[K,J])    -- depth B*T in one pass
//
// The lead rounds are rounds of classic blocked FW restricted to the cross
// region (rows or columns in K), which is closed under those updates.  The
// rest pass uses the stage's final panels: for i, j outside K every path
// through a vertex of K splits at it into two paths whose lengths the final
// panels bound, so one product of depth B*T gives the same values as B
// rounds.  Each rest tile therefore makes one HBM round trip per kappa
// k-tiles (the point of the reference's multi-layer scheme), with C held in
// VGPRs across the whole depth.  kappa = 1 is the classic three-phase
// algorithm, which is what tiling_gpu runs.
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <mutex>
#include <vector>

#include "fwx_device.hpp"
#include "fwx_internal.hpp"
#include "fwx_tile.hpp"
#include "fwx_tile16.hpp"

namespace fwx {

constexpr int kChunk = 32;
constexpr int kDiagThreads = 512;

// The lead kernels (diag, panel, cross) form the look-ahead schedule's
// dependent chain and share CUs with the concurrent rest pass; raising their
// waves' issue priority lets them take the SIMDs first where both are
// resident.  A wave's own priority, set by the wave; it ends with the wave.
__device__ __forceinline__ void lead_wave_priority() { __builtin_amdgcn_s_setprio(3); }

// ---------------------------------------------------------------------------
// diag: close tile (c, c).
template <int T, int NTH>
__global__ __launch_bounds__(NTH) void k_diag(int32_t* __restrict__ D, int32_t n, int32_t c, int32_t prio) {
  if (prio) lead_wave_priority();
  close_tile<T, NTH>(make_view(D, n, c * T, c * T));
}

// diag closed by pk16 squaring when every cell of the tile is below
// kSq16Small and the tile has no padding (close_sq16_regs), else by the T
// steps; same closure either way.
template <int T, int NTH>
__global__ __launch_bounds__(NTH) void k_diag_sq(int32_t* __restrict__ D, int32_t n, int32_t c, int32_t prio) {
  if (prio) lead_wave_priority();
  __shared__ SqSmem16<T, NTH> s;
  close_tile_sq<T, NTH>(make_view(D, n, c * T, c * T), s);
}

// ---------------------------------------------------------------------------
// panel: blocks [0, N-1) take row tile (c, j), blocks [N-1, 2N-2) column tile
// (i, c).  The tile's own pre-update values are the second operand; it is
// staged chunk by chunk before this block stores anything.
template <int T, bool VEC, bool PK>
__global__ __launch_bounds__(256) void k_panel(int32_t* __restrict__ D, int32_t n, int32_t c,
                                               int32_t N, int32_t prio) {
  if (prio) lead_wave_priority();
  __shared__ TileSmem<T, T, kChunk> s;
  const int b = blockIdx.x;
  const int32_t base = c * T;
  const View diag = make_view(D, n, base, base);
  View A, B, C;
  if (b < N - 1) {
    const int j = b + (b >= c);
    C = make_view(D, n, base, j * T);
    A = diag;
    B = C;
  } else {
    int i = b - (N - 1);
    i += (i >= c);
    C = make_view(D, n, i * T, base);
    A = C;
    B = diag;
  }
  tile_update<T, T, kChunk, VEC, 16, PK>(s, A, B, C, T);
}

// panel in the pk16 form (64-wide tiles, 256 threads of 4 x 4 cells).
template <int T, bool VEC>
__global__ __launch_bounds__(256) void k_panel16(int32_t* __restrict__ D, int32_t n, int32_t c, int32_t N,
                                                 int32_t prio) {
  if (prio) lead_wave_priority();
  __shared__ TileSmem16<T, T, 16> s;
  const int b = blockIdx.x;
  const int32_t base = c * T;
  const View diag = make_view(D, n, base, base);
  View A, B, C;
  if (b < N - 1) {
    const int j = b + (b >= c);
    C = make_view(D, n, base, j * T);
    A = diag;
    B = C;
  } else {
    int i = b - (N - 1);
    i += (i >= c);
    C = make_view(D, n, i * T, base);
    A = C;
    B = diag;
  }
  tile_update16<T, T, 16, VEC>(s, A, B, C, T);
}

// ---------------------------------------------------------------------------
// dpanel: diag and panel fused into one launch.  Every block first closes the
// diagonal tile itself and stores it (all blocks store the same bits), then
// updates its row/column tile as k_panel does, reading the closed tile back.
// A block may start from a partly overwritten diagonal tile; every cell then
// lies between the closed value and the original one, and closing such a
// tile yields the same closed tile, so the stored bits never change.
// Saves one dependent launch per k-tile round at the price of 2(N-1)
// redundant T^3 closures (negligible next to the round's products).
template <int T, bool VEC, bool PK>
__global__ __launch_bounds__(256) void k_dpanel(int32_t* __restrict__ D, int32_t n, int32_t c,
                                                int32_t N, int32_t prio) {
  if (prio) lead_wave_priority();
  close_tile<T>(make_view(D, n, c * T, c * T));
  __syncthreads();
  __shared__ TileSmem<T, T, kChunk> s;
  const int b = blockIdx.x;
  const int32_t base = c * T;
  const View diag = make_view(D, n, base, base);
  View A, B, C;
  if (b < N - 1) {
    const int j = b + (b >= c);
    C = make_view(D, n, base, j * T);
    A = diag;
    B = C;
  } else {
    int i = b - (N - 1);
    i += (i >= c);
    C = make_view(D, n, i * T, base);
    A = C;
    B = diag;
  }
  tile_update<T, T, kChunk, VEC, 16, PK>(s, A, B, C, T);
}

// ---------------------------------------------------------------------------
// cross: tiles with a row or a column in the stage, other than row c and
// column c.  Region 1: stage rows (except c) x all columns (except c),
// (B-1)(N-1) tiles.  Region 2: non-stage rows x stage columns (except c),
// (N-B)(B-1) tiles.
template <int T, bool VEC, bool PK>
__global__ __launch_bounds__(256) void k_cross(int32_t* __restrict__ D, int32_t n, int32_t c,
                                               int32_t kb, int32_t Bst, int32_t N, int32_t prio) {
  if (prio) lead_wave_priority();
  __shared__ TileSmem<T, T, kChunk> s;
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
  const View A = make_view(D, n, i * T, c * T);
  const View B = make_view(D, n, c * T, j * T);
  const View C = make_view(D, n, i * T, j * T);
  tile_update<T, T, kChunk, VEC, 16, PK>(s, A, B, C, T);
}

// cross in the pk16 form (64-wide tiles).
// Round c's update of tile (c+1, c+1) (a cross tile when c+1 lies in the
// stage) followed by its closure -- the next round's diagonal -- in the
// block that owns the tile, so the next round needs no diagonal launch.  No
// other block of the launch reads or writes that tile.  The product's
// result goes through LDS into the closure's register layout: a cell whose
// pk16 result reached the cap keeps its input value (as store_c16 does).
template <int T, bool VEC, bool EDGE>
__device__ __forceinline__ void cross_and_close(TileSmem16<T, T, 16>& ps, SqSmem16<T, 256>& ss, const View& A,
                                                const View& B, const View& C) {
  using P = TileProduct16<T, T, 16, VEC, EDGE>;
  uint32_t acc[P::TM][P::TP];
  P::load_c16(C, acc);
  P p;
  p.run16(ps, A, B, T, acc);  // ends with a barrier: the product's LDS is free
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  int32_t* tile = reinterpret_cast<int32_t*>(ss.a);
#pragma unroll
  for (int r = 0; r < P::TM; ++r)
#pragma unroll
    for (int g = 0; g < P::TG; ++g) {
      const int i = ty + 16 * r, j = 4 * (tx + 16 * g);
      const uint32_t lo = acc[r][2 * g] - kP16Bias2, hi = acc[r][2 * g + 1] - kP16Bias2;
      const uint32_t e[4] = {lo & 0xFFFFu, lo >> 16, hi & 0xFFFFu, hi >> 16};
#pragma unroll
      for (int x = 0; x < 4; ++x)
        tile[i * T + j + x] = e[x] < kP16Cap ? (int32_t)e[x] : (C.in(i, j + x) ? *C.at(i, j + x) : kInf);
    }
  __syncthreads();
  constexpr int TR = T / 16, TQ = T / 16;
  int32_t cl[TR][TQ];
#pragma unroll
  for (int r = 0; r < TR; ++r)
#pragma unroll
    for (int q = 0; q < TQ; ++q) cl[r][q] = tile[(ty + 16 * r) * T + tx + 16 * q];
  __syncthreads();  // the closure reuses the buffer
  const int kmax = min(T, C.rlim);
  if (!close_sq16_regs<T, 256>(cl, ss, kmax)) close_regs<T, 256>(cl, reinterpret_cast<int32_t*>(ss.a), kmax);
#pragma unroll
  for (int r = 0; r < TR; ++r)
#pragma unroll
    for (int q = 0; q < TQ; ++q)
      if (C.in(ty + 16 * r, tx + 16 * q)) *C.at(ty + 16 * r, tx + 16 * q) = cl[r][q];
}

template <int T, bool VEC>
__global__ __launch_bounds__(256) void k_cross16(int32_t* __restrict__ D, int32_t n, int32_t c, int32_t kb,
                                                 int32_t Bst, int32_t N, int32_t prio, int32_t close_next) {
  if (prio) lead_wave_priority();
  __shared__ union {
    TileSmem16<T, T, 16> prod;
    SqSmem16<T, 256> sq;
  } s;
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
  const View A = make_view(D, n, i * T, c * T);
  const View B = make_view(D, n, c * T, j * T);
  const View C = make_view(D, n, i * T, j * T);
  // block l (N - 1) + c holds tile (c+1, c+1)
  if (close_next && (int)blockIdx.x == l * (N - 1) + c) {
    if (C.rlim >= T && C.clim >= T)
      cross_and_close<T, VEC, false>(s.prod, s.sq, A, B, C);
    else
      cross_and_close<T, VEC, true>(s.prod, s.sq, A, B, C);
    return;
  }
  tile_update16<T, T, 16, VEC>(s.prod, A, B, C, T);
}

// ---------------------------------------------------------------------------
// Range gate of the kernels of a packed pk16 solve (k_block_square16,
// k_block_panel16[w], k_pack16, k_rest16p, k_rest_band16[w]).  Such a solve
// enqueues its range check without waiting for it; `gate` is the word that
// check ORs into, and a kernel returns at entry, writing nothing, when an
// input cell lies outside [0, 2^30).  The host reads the word after the
// solve and, if it is set, solves the untouched input in the exact form.
// Null in every other solve.  The check completed before any gated kernel
// starts (stream and event order), so the word is final when it is read.
__device__ __forceinline__ bool gate_closed(const int* gate) { return gate != nullptr && *gate != 0; }

// ---------------------------------------------------------------------------
// Block lead phase (pk16, 64-wide tiles): the stage's kappa lead rounds as
// the three-phase algorithm with block size L = B*T.
//
// block_square: the L x L stage block closes by min-plus squaring, block <-
// min(block, block (x) block), one launch per squaring: workgroup (a, b)
// updates tile (kb+a, kb+b) in place with one product of depth L.  A tile
// reads a mix of the other tiles' old and new values; every value stays a
// walk length at least the closure's and at most the synchronous squaring's,
// so log2(L) launches reach the closure, and a launch in which no cell
// changed leaves the block obeying every triangle inequality: the closure.
// A workgroup that changed a cell writes the stage's generation `gen` into
// flags[it]; launch it returns at once unless flags[it-1] holds gen, so the
// launches after convergence cost only their dispatch.  (A stale flag equal
// to gen can only add a squaring; it cannot stop one early.)  In the pk16
// form a cell whose result reaches the cap keeps its value, as in every pk16
// product: the block is exact below the cap and the solve's final check
// covers the rest.
template <int T, bool VEC>
__global__ __launch_bounds__(256) void k_block_square16(int32_t* __restrict__ D, int32_t n, int32_t kb, int32_t B,
                                                        int32_t it, uint32_t* __restrict__ flags, uint32_t gen,
                                                        int32_t prio, const int* __restrict__ gate) {
  if (gate_closed(gate) || (it > 0 && flags[it - 1] != gen)) return;
  if (prio) lead_wave_priority();
  __shared__ TileSmem16<T, T, 16> s;
  const int a = kb + (int)blockIdx.x / B, b = kb + (int)blockIdx.x % B;
  const int32_t s0 = kb * T;
  const View A = make_view(D, n, a * T, s0);
  const View Bv = make_view(D, n, s0, b * T);
  const View C = make_view(D, n, a * T, b * T);
  if (tile_update16_changed<T, T, 16, VEC>(s, A, Bv, C, B * T) && threadIdx.x == 0) flags[it] = gen;
}

// block_square with each tile's depth split over KS workgroups (grid
// B B KS, a tile's parts in adjacent workgroups): part q takes k in
// [q L / KS, (q + 1) L / KS) and lowers the tile's changed cells by atomic
// min (tile_update16_part).  Every value written is still a walk length at
// least the closure's, and after the launch each cell is at most the
// minimum over all k of the values read, hence at most the synchronous
// squaring's: the launch count and the early exit of k_block_square16 hold
// unchanged.  KS times the workgroups of depth L / KS, for a stage block
// whose lead phase has the GPU to itself (the first stage).  T = 64, full
// tiles (n % 64 == 0).
template <int KS, bool VEC>
__global__ __launch_bounds__(256) void k_block_square16s(int32_t* __restrict__ D, int32_t n, int32_t kb, int32_t B,
                                                         int32_t it, uint32_t* __restrict__ flags, uint32_t gen,
                                                         int32_t prio, const int* __restrict__ gate) {
  if (gate_closed(gate) || (it > 0 && flags[it - 1] != gen)) return;
  if (prio) lead_wave_priority();
  constexpr int T = 64;
  __shared__ TileSmem16<T, T, 16> s;
  const int t = (int)blockIdx.x / KS, q = (int)blockIdx.x % KS;
  const int a = kb + t / B, b = kb + t % B;
  const int32_t part = B * T / KS;
  const int32_t k0 = kb * T + q * part;
  const View A = make_view(D, n, a * T, k0);
  const View Bv = make_view(D, n, k0, b * T);
  const View C = make_view(D, n, a * T, b * T);
  if (tile_update16_part<T, T, 16, VEC>(s, A, Bv, C, part) && threadIdx.x == 0) flags[it] = gen;
}

// block_panel: every row-panel tile (k, j) and column-panel tile (i, k) of
// the stage (k in the stage, i, j outside) takes one product of depth L with
// the closed block: D[k, j] <- min(D[k, j], D*[k, K] (x) D[K, j]) and
// D[i, k] <- min(D[i, k], D[i, K] (x) D*[K, k]).  The other operand is the
// panel being updated in the same launch; a mix of its old and new values
// gives the same result (the new values lie between the final panel and the
// old one, and the final panel is closed under the block).  Blocks
// [0, B (N-B)) take the row tiles, the rest the column tiles.
template <int T, bool VEC>
__global__ __launch_bounds__(256) void k_block_panel16(int32_t* __restrict__ D, int32_t n, int32_t kb, int32_t B,
                                                       int32_t N, int32_t prio, const int* __restrict__ gate) {
  if (gate_closed(gate)) return;
  if (prio) lead_wave_priority();
  __shared__ TileSmem16<T, T, 16> s;
  const int r = B * (N - B);
  int b = blockIdx.x, i, j;
  if (b < r) {
    i = kb + b / (N - B);
    const int jj = b % (N - B);
    j = jj + (jj >= kb ? B : 0);
  } else {
    b -= r;
    const int ii = b / B;
    i = ii + (ii >= kb ? B : 0);
    j = kb + b % B;
  }
  const int32_t s0 = kb * T;
  const View A = make_view(D, n, i * T, s0);
  const View Bv = make_view(D, n, s0, j * T);
  const View C = make_view(D, n, i * T, j * T);
  tile_update16<T, T, 16, VEC>(s, A, Bv, C, B * T);
}

// block_panel on 64 x 128 tiles (two adjacent 64-wide panel tiles per
// workgroup, 4 x 8 cells per thread: a quarter fewer LDS reads per
// relaxation than 64 x 64).  Needs B and kb even, so that the paired columns
// are adjacent: row-panel pairs (jj, jj + 1) of the columns outside the
// stage, column-panel pairs of stage columns.  Blocks [0, B (N-B) / 2) take
// the row pairs.
template <bool VEC>
__global__ __launch_bounds__(256) void k_block_panel16w(int32_t* __restrict__ D, int32_t n, int32_t kb, int32_t B,
                                                        int32_t N, int32_t prio, const int* __restrict__ gate) {
  if (gate_closed(gate)) return;
  if (prio) lead_wave_priority();
  constexpr int T = 64;
  __shared__ TileSmem16<T, 2 * T, 16> s;
  const int h = (N - B) / 2;
  const int r = B * h;
  int b = blockIdx.x, i, j;
  if (b < r) {
    i = kb + b / h;
    const int jj = 2 * (b % h);
    j = jj + (jj >= kb ? B : 0);
  } else {
    b -= r;
    const int ii = b / (B / 2);
    i = ii + (ii >= kb ? B : 0);
    j = kb + 2 * (b % (B / 2));
  }
  const int32_t s0 = kb * T;
  const View A = make_view(D, n, i * T, s0);
  const View Bv = make_view(D, n, s0, j * T);
  const View C = make_view(D, n, i * T, j * T);
  tile_update16<T, 2 * T, 16, VEC>(s, A, Bv, C, B * T);
}

// ---------------------------------------------------------------------------
// rest: R x R tiles outside the stage rows/columns, depth S = B*T.
// The (m x m) reduced tile grid is split into 8 contiguous ranges of `per`
// tiles, one per XCD (workgroups are dealt to the 8 XCDs round-robin), and
// each range is walked in 8-row groups so concurrently resident tiles share
// A panels (same rows) and B panels (nearby columns) in that XCD's L2.
__device__ __forceinline__ bool rest_tile_of_block(int m, int per, int sr0, int srn, int* I, int* J) {
  const int b = blockIdx.x;
  const int L = (b & 7) * per + (b >> 3);
  if (L >= m * m) return false;
  constexpr int G = 8;
  const int group = L / (G * m);
  const int first = group * G;
  const int rows = min(m - first, G);
  const int within = L - group * G * m;
  const int Ii = first + within % rows;
  const int Jj = within / rows;
  *I = Ii + (Ii >= sr0 ? srn : 0);
  *J = Jj + (Jj >= sr0 ? srn : 0);
  return true;
}

template <int R, bool VEC, bool PK, int KC = kChunk>
__global__ __launch_bounds__(256, KC == 16 ? 3 : 2) void k_rest(int32_t* __restrict__ D, int32_t n,
                                                                        int32_t s0, int32_t S, int32_t sr0,
                                                                        int32_t srn, int32_t m, int32_t per) {
  __shared__ TileSmem<R, R, KC> s;
  int I, J;
  if (!rest_tile_of_block(m, per, sr0, srn, &I, &J)) return;
  const View A = make_view(D, n, I * R, s0);
  const View B = make_view(D, n, s0, J * R);
  const View C = make_view(D, n, I * R, J * R);
  tile_update<R, R, KC, VEC, 16, PK>(s, A, B, C, S);
}

// 512-thread shape, held to 128 VGPRs so that two blocks (4 waves/SIMD)
// share a CU.
template <int R, bool VEC, bool PK>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4, 4))) void k_rest512(
    int32_t* __restrict__ D, int32_t n, int32_t s0, int32_t S, int32_t sr0, int32_t srn, int32_t m,
    int32_t per) {
  __shared__ TileSmem<R, R, kChunk> s;
  int I, J;
  if (!rest_tile_of_block(m, per, sr0, srn, &I, &J)) return;
  const View A = make_view(D, n, I * R, s0);
  const View B = make_view(D, n, s0, J * R);
  const View C = make_view(D, n, I * R, J * R);
  tile_update<R, R, kChunk, VEC, 32, PK>(s, A, B, C, S);
}

// pk16 shape (fwx_tile16.hpp): 256 threads, 8 x 8 cells each in 32 packed
// VGPRs.  KC = 32: 52 KB of LDS, three blocks per CU; KC = 16: 26 KB, four
// (held to 128 VGPRs).
template <int R, bool VEC, int KC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(KC == 16 ? 4 : 3, 4))) void k_rest16(
    int32_t* __restrict__ D, int32_t n, int32_t s0, int32_t S, int32_t sr0, int32_t srn, int32_t m,
    int32_t per) {
  __shared__ TileSmem16<R, R, KC> s;
  int I, J;
  if (!rest_tile_of_block(m, per, sr0, srn, &I, &J)) return;
  const View A = make_view(D, n, I * R, s0);
  const View B = make_view(D, n, s0, J * R);
  const View C = make_view(D, n, I * R, J * R);
  tile_update16<R, R, KC, VEC>(s, A, B, C, S);
}

// Stage panels packed for k_rest16p (PackedPanels in fwx_tile16.hpp): blocks
// [0, na) write wa[i * S + k] = p16_dup(D[i][s0 + k]) (i < n, k < S; one
// 16-byte group of k per thread), the others wb[p * (n / 2) + c] =
// p16_pair(D[s0 + 2p][2c], D[s0 + 2p][2c + 1]) | p16_pair(row s0 + 2p + 1)
// << 32 (p < S / 2; two slots per thread).  Needs n % 4 == 0, S % 4 == 0,
// s0 + S <= n, a 16-byte aligned D and cells in [0, 2^30) (the pk16 form's
// precondition).
__global__ __launch_bounds__(256) void k_pack16(const int32_t* __restrict__ D, int32_t n, int32_t s0, int32_t S,
                                                uint32_t* __restrict__ wa, uint64_t* __restrict__ wb,
                                                int32_t na, int32_t prio, const int* __restrict__ gate,
                                                int* __restrict__ capped) {
  if (gate_closed(gate)) return;
  if (prio) lead_wave_priority();
  int32_t mx = 0;
  if ((int32_t)blockIdx.x < na) {
    const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t i = q / (S / 4);
    const int32_t k = (int32_t)(q % (S / 4)) * 4;
    if (i < n) {
      const int4 v = *reinterpret_cast<const int4*>(D + i * n + s0 + k);
      *reinterpret_cast<uint4*>(wa + i * S + k) =
          make_uint4(p16_dup(v.x), p16_dup(v.y), p16_dup(v.z), p16_dup(v.w));
      mx = max(max(v.x, v.y), max(v.z, v.w));
    }
  } else {
    const int64_t q = (int64_t)(blockIdx.x - na) * 256 + threadIdx.x;
    const int64_t p = q / (n / 4);
    const int32_t c4 = (int32_t)(q % (n / 4)) * 4;
    if (p < S / 2) {
      const int4 r0 = *reinterpret_cast<const int4*>(D + (s0 + 2 * p) * (int64_t)n + c4);
      const int4 r1 = *reinterpret_cast<const int4*>(D + (s0 + 2 * p + 1) * (int64_t)n + c4);
      ulonglong2 v;
      v.x = (uint64_t)p16_pair(r0.x, r0.y) | ((uint64_t)p16_pair(r1.x, r1.y) << 32);
      v.y = (uint64_t)p16_pair(r0.z, r0.w) | ((uint64_t)p16_pair(r1.z, r1.w) << 32);
      *reinterpret_cast<ulonglong2*>(wb + p * (n / 2) + c4 / 2) = v;
      mx = max(max(max(r0.x, r0.y), max(r0.z, r0.w)), max(max(r1.x, r1.y), max(r1.z, r1.w)));
    }
  }
  if (capped && __ballot(mx >= (int32_t)kP16Cap) != 0ull && (threadIdx.x & 63) == 0) atomicOr(capped, 1);
}

// The pk16 rest pass on packed panels: the tiles and XCD mapping of
// k_rest16, 16-deep chunks staged by plain copies.  n % R == 0 (every tile
// whole), S % 16 == 0.
template <int R, int WAVES>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WAVES, WAVES))) void k_rest16p(
    int32_t* __restrict__ D, int32_t n, const uint32_t* __restrict__ wa, const uint64_t* __restrict__ wb, int32_t S,
    int32_t sr0, int32_t srn, int32_t m, int32_t per, const int* __restrict__ gate, int* __restrict__ capped) {
  __shared__ TileSmem16<R, R, 16> s;
  int I, J;
  if (gate_closed(gate) || !rest_tile_of_block(m, per, sr0, srn, &I, &J)) return;
  const PackedPanels w{wa + (int64_t)I * R * S, S, wb + (int64_t)J * (R / 2), n / 2};
  const View C = make_view(D, n, I * R, J * R);
  tile_update16_packed<R, R, 16>(s, w, C, S, capped);
}

// ---------------------------------------------------------------------------
// rest_band: the part of stage s's rest pass that the next stage's lead rounds
// read -- R x R tiles outside stage s's band [e0, e0+en) with a row or a
// column in the next stage's band [f0, f0+fn) (R-tile units, f0 = e0 + en).
// m1 = tiles per side outside band e.  Region 1: band-f rows x every non-e
// column (fn * m1 tiles); region 2: rows outside both bands x band-f columns.
template <int R, bool VEC, bool PK>
__global__ __launch_bounds__(256) void k_rest_band(int32_t* __restrict__ D, int32_t n, int32_t s0,
                                                   int32_t S, int32_t e0, int32_t en, int32_t f0,
                                                   int32_t fn, int32_t m1) {
  __shared__ TileSmem<R, R, kChunk> s;
  int b = blockIdx.x;
  int I, J;
  if (b < fn * m1) {
    I = f0 + b / m1;
    const int Jr = b % m1;
    J = Jr + (Jr >= e0 ? en : 0);
  } else {
    b -= fn * m1;
    const int Ir = b / fn;  // among rows outside both bands
    I = Ir + (Ir >= e0 ? en + fn : 0);
    J = f0 + b % fn;
  }
  const View A = make_view(D, n, I * R, s0);
  const View B = make_view(D, n, s0, J * R);
  const View C = make_view(D, n, I * R, J * R);
  tile_update<R, R, kChunk, VEC, 16, PK>(s, A, B, C, S);
}

// rest_band in the pk16 form (64-wide tiles: 256 threads of 4 x 4 cells,
// 16-deep chunks); the band is part of stage s's rest pass, so its cells
// follow the same rule as the pk16 rest passes (a cell at the cap keeps its
// value; the solve's check afterwards finds it).
template <int R, bool VEC>
__global__ __launch_bounds__(256) void k_rest_band16(int32_t* __restrict__ D, int32_t n, int32_t s0, int32_t S,
                                                     int32_t e0, int32_t en, int32_t f0, int32_t fn, int32_t m1,
                                                     const int* __restrict__ gate) {
  if (gate_closed(gate)) return;
  __shared__ TileSmem16<R, R, 16> s;
  int b = blockIdx.x;
  int I, J;
  if (b < fn * m1) {
    I = f0 + b / m1;
    const int Jr = b % m1;
    J = Jr + (Jr >= e0 ? en : 0);
  } else {
    b -= fn * m1;
    const int Ir = b / fn;
    I = Ir + (Ir >= e0 ? en + fn : 0);
    J = f0 + b % fn;
  }
  const View A = make_view(D, n, I * R, s0);
  const View B = make_view(D, n, s0, J * R);
  const View C = make_view(D, n, I * R, J * R);
  tile_update16<R, R, 16, VEC>(s, A, B, C, S);
}

// rest_band16 on 64 x 128 tiles: two adjacent 64-wide band tiles per
// workgroup (4 x 8 cells per thread).  Needs e0, en, f0, fn and m1 even, so
// that the paired columns are adjacent.  Blocks [0, fn m1 / 2) take band-f
// rows x column pairs outside band e; the others rows outside both bands x
// band-f column pairs.
template <bool VEC>
__global__ __launch_bounds__(256) void k_rest_band16w(int32_t* __restrict__ D, int32_t n, int32_t s0, int32_t S,
                                                      int32_t e0, int32_t en, int32_t f0, int32_t fn, int32_t m1,
                                                      int32_t prio, const int* __restrict__ gate) {
  if (gate_closed(gate)) return;
  if (prio) lead_wave_priority();
  constexpr int R = 64;
  __shared__ TileSmem16<R, 2 * R, 16> s;
  int b = blockIdx.x;
  int I, J;
  const int r1 = fn * (m1 / 2);
  if (b < r1) {
    I = f0 + b / (m1 / 2);
    const int Jr = 2 * (b % (m1 / 2));
    J = Jr + (Jr >= e0 ? en : 0);
  } else {
    b -= r1;
    const int Ir = b / (fn / 2);
    I = Ir + (Ir >= e0 ? en + fn : 0);
    J = f0 + 2 * (b % (fn / 2));
  }
  const View A = make_view(D, n, I * R, s0);
  const View B = make_view(D, n, s0, J * R);
  const View C = make_view(D, n, I * R, J * R);
  tile_update16<R, 2 * R, 16, VEC>(s, A, B, C, S);
}

// ---------------------------------------------------------------------------
// host driver

// FWX_PANEL_WIDE=0 / FWX_BAND_WIDE=0 keep the block panels / the pk16 band
// and X passes on 64 x 64 tiles instead of 64 x 128 (measured at n=8192 on
// (64, 8): block panels on 64 x 128 tiles 9.47 -> 9.36 ms per step, then
// band and X passes on 64 x 128 tiles 9.32 -> 9.25 ms).
static bool panel_wide() {
  const char* e = std::getenv("FWX_PANEL_WIDE");
  return !(e && e[0] == '0');
}
static bool band_wide() {
  const char* e = std::getenv("FWX_BAND_WIDE");
  return !(e && e[0] == '0');
}

// Packed stage panels for the pk16 rest passes (k_pack16 -> k_rest16p).  A
// solve holds one workspace of three panel buffers: stage s packs into buffer
// s % 3 after its lead rounds, on the stream that ran them, and its rest pass
// reads that buffer.  Three, because in the side-band-ahead schedule the side
// stream packs stage s + 1 while the caller's stream may still run the rest
// pass of stage s - 1 (band(s) waits only for X(s - 1), which the caller's
// stream runs after Y(s - 2)); stage s + 1's buffer is then stage s - 2's,
// whose rest pass has ended.  The other schedules order packs after older
// rest passes too.  Workspaces are allocated on first need per device, kept
// for reuse and never freed; a solve that cannot get one runs unpacked.
constexpr int kMaxSquarings = 16;  // log2 of the widest stage block (16 x 128) is 11

struct PackWs {
  int dev = -1;
  size_t cap_a = 0, cap_b = 0;  // words / slots per buffer
  uint32_t* wa[3] = {};
  uint64_t* wb[3] = {};
  uint32_t* flags = nullptr;  // kMaxSquarings words: the block squarings' change flags
  uint32_t gen = 0;           // generation of the last stage that used them
  int* gate = nullptr;        // two words after the flags: the deferred range and cap checks
  bool busy = false;
  // the holding solve's configuration
  int32_t tile = 0, kappa = 0;
};

static std::mutex g_pack_mu;
static std::vector<PackWs*> g_pack;
static thread_local PackWs* t_pack = nullptr;
// The range gate of the solve being enqueued (its workspace's gate word), or
// null when the range check has already been read.
static thread_local const int* t_gate = nullptr;
// The word that the last stage's pack and rest pass OR "a final cell at the
// pk16 cap" into (the gate's second word), or null: then a separate pass
// over the matrix checks it after the solve.
static thread_local int* t_capped = nullptr;

// A launch site without the range gate, reached while one is set: the
// eligibility test of the deferred check (range_deferrable) admitted a
// configuration that leaves the gated kernels.  Refused rather than run on
// an input that may be outside the form's range.
static int ungated_launch(const char* what) {
  return fail(FWX_ERR_UNSUPPORTED, std::string("internal: ") + what + " launched in a solve with a deferred range check");
}

// FWX_PACK16=0 runs the pk16 rest passes on the int32 panels (conversion in
// every block).
static bool pack16_enabled() {
  const char* e = std::getenv("FWX_PACK16");
  return !(e && e[0] == '0');
}

// Waves per SIMD of k_rest16p: 5 (its 90 VGPRs and 24 KB of LDS fit five
// blocks per CU; the int32-panel kernel's 116 VGPRs held it to four), or 4
// with FWX_REST16P_WAVES=4.  Measured on MI355X: n=8192 (64, 4) rest launch
// 265.9 vs 269.1 us, alone 279.6 vs 283.0 us; n=32768 (64, 8) 0.528 vs
// 0.536 s per solve (int32 panels: 271.6 / 291.8 us, 0.547 s).
static int rest16p_waves() {
  const char* e = std::getenv("FWX_REST16P_WAVES");
  return (e && e[0] == '4') ? 4 : 5;
}

static PackWs* pack_acquire(int32_t n, int32_t tile, int32_t kappa) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  const int64_t span = (int64_t)tile * kappa;
  const size_t need_a = (size_t)n * (size_t)span, need_b = (size_t)(span / 2) * (size_t)(n / 2);
  std::lock_guard<std::mutex> lk(g_pack_mu);
  PackWs* w = nullptr;
  for (PackWs* c : g_pack)
    if (!c->busy && c->dev == dev && c->cap_a >= need_a && c->cap_b >= need_b) {
      w = c;
      break;
    }
  if (!w) {
    char* base = nullptr;
    if (hipMalloc(reinterpret_cast<void**>(&base), 3 * (need_a * 4 + need_b * 8) + 256) != hipSuccess) {
      (void)hipGetLastError();
      return nullptr;
    }
    w = new PackWs;
    w->dev = dev;
    w->cap_a = need_a;
    w->cap_b = need_b;
    for (int i = 0; i < 3; ++i) {
      w->wb[i] = reinterpret_cast<uint64_t*>(base + i * need_b * 8);
      w->wa[i] = reinterpret_cast<uint32_t*>(base + 3 * need_b * 8 + i * need_a * 4);
    }
    w->flags = reinterpret_cast<uint32_t*>(base + 3 * (need_a * 4 + need_b * 8));
    w->gate = reinterpret_cast<int*>(w->flags + kMaxSquarings);
    g_pack.push_back(w);
  }
  w->busy = true;
  w->tile = tile;
  w->kappa = kappa;
  return w;
}

static void pack_release(PackWs* w) {
  if (!w) return;
  std::lock_guard<std::mutex> lk(g_pack_mu);
  w->busy = false;
}

// Packs stage kb's panels (s0 = kb T, width S) into its buffer, on st.
static int pack_stage(int32_t* d, int32_t n, int32_t kb, int32_t S, int prio, hipStream_t st) {
  const PackWs* w = t_pack;
  const int b = (kb / w->kappa) % 3;
  const int32_t s0 = kb * w->tile;
  const int64_t ga = (int64_t)n * (S / 4), gb = (int64_t)(S / 2) * (n / 4);
  const int na = (int)((ga + 255) / 256), nb = (int)((gb + 255) / 256);
  // the last stage's panels are final when packed: every cell with a row or
  // a column in the stage
  int* capped = t_capped && s0 + S >= n ? t_capped : nullptr;
  k_pack16<<<na + nb, 256, 0, st>>>(d, n, s0, S, w->wa[b], w->wb[b], na, prio, t_gate, capped);
  FWX_LAUNCHED("k_pack16");
  return FWX_OK;
}

// FWX_DPANEL=1 fuses diag and panel into one launch (measured slower on
// MI355X: the redundant closures contend with the rest pass).
static bool fused_dpanel() {
  const char* e = std::getenv("FWX_DPANEL");
  return e && e[0] == '1';
}

// Threads of the diagonal-closure block (FWX_DIAG_NTH = 256 / 512 / 1024):
// more threads hold fewer cells each and give more waves per SIMD to hide
// the latency of the T dependent steps.  512 measured best or tied on MI355X
// (n=1024, tile 64: 0.455 ms per solve vs 0.510 ms with 256; n=4096, tile
// 128: 3.83 vs 3.96 ms).
static int diag_threads() {
  const char* e = std::getenv("FWX_DIAG_NTH");
  const int v = e ? std::atoi(e) : 0;
  return (v == 256 || v == 512 || v == 1024) ? v : kDiagThreads;
}

// FWX_DIAG_SQ=0 closes every diagonal tile by the T Floyd-Warshall steps;
// else 32- and 64-wide tiles whose cells are all below kSq16Small close by
// pk16 squaring (k_diag_sq).
static bool diag_sq_enabled() {
  const char* e = std::getenv("FWX_DIAG_SQ");
  return !(e && e[0] == '0');
}

template <int T>
static int diag_launch(int32_t* d, int32_t n, int32_t c, int prio, hipStream_t st) {
  if (T <= 64 && diag_sq_enabled()) {
    if (T == 32)
      k_diag_sq<32, 256><<<1, 256, 0, st>>>(d, n, c, prio);
    else
      k_diag_sq<64, 512><<<1, 512, 0, st>>>(d, n, c, prio);
    FWX_LAUNCHED("k_diag_sq");
    return FWX_OK;
  }
  const int nth = T == 32 ? 256 : diag_threads();
  if (nth == 1024)
    k_diag<T, T == 32 ? 256 : 1024><<<1, T == 32 ? 256 : 1024, 0, st>>>(d, n, c, prio);
  else if (nth == 512)
    k_diag<T, T == 32 ? 256 : 512><<<1, T == 32 ? 256 : 512, 0, st>>>(d, n, c, prio);
  else
    k_diag<T, 256><<<1, 256, 0, st>>>(d, n, c, prio);
  FWX_LAUNCHED("k_diag");
  return FWX_OK;
}

// FWX_LEAD16=0 keeps the 64-wide lead products in pk64 when the solve runs
// pk16 rest passes.
static bool lead16_enabled() {
  const char* e = std::getenv("FWX_LEAD16");
  return !(e && e[0] == '0');
}

// FWX_LEAD_CLOSE=0 gives every pk16 lead round its own diagonal launch;
// else round c's cross launch also closes tile (c+1, c+1) (cross_and_close)
// and only a stage's first round launches k_diag.
static bool lead_close_enabled() {
  const char* e = std::getenv("FWX_LEAD_CLOSE");
  return !(e && e[0] == '0');
}

template <int T, bool VEC, bool PK, bool P16 = false>
static int lead_round(int32_t* d, int32_t n, int32_t c, int32_t kb, int32_t B, int32_t N, int prio,
                      hipStream_t st) {
  if (t_gate) return ungated_launch("lead round");
  if (P16 && T == 64 && lead16_enabled()) {
    // with B > 1 every round but the stage's last has a cross launch, which
    // then closes the next round's diagonal tile
    const int l = c - kb;
    const bool chain = lead_close_enabled() && B > 1;
    if (l == 0 || !chain)
      if (int rc = diag_launch<T>(d, n, c, prio, st)) return rc;
    if (N > 1) {
      k_panel16<64, VEC><<<2 * (N - 1), 256, 0, st>>>(d, n, c, N, prio);
      FWX_LAUNCHED("k_panel16");
    }
    if (B > 1) {
      const int cnt = (B - 1) * (2 * N - B - 1);
      if (cnt > 0) {
        k_cross16<64, VEC><<<cnt, 256, 0, st>>>(d, n, c, kb, B, N, prio, chain && l + 1 < B ? 1 : 0);
        FWX_LAUNCHED("k_cross16");
      }
    }
    return FWX_OK;
  }
  if (N > 1 && fused_dpanel()) {
    k_dpanel<T, VEC, PK><<<2 * (N - 1), 256, 0, st>>>(d, n, c, N, prio);
    FWX_LAUNCHED("k_dpanel");
  } else {
    if (int rc = diag_launch<T>(d, n, c, prio, st)) return rc;
    if (N > 1) {
      k_panel<T, VEC, PK><<<2 * (N - 1), 256, 0, st>>>(d, n, c, N, prio);
      FWX_LAUNCHED("k_panel");
    }
  }
  if (B > 1) {
    const int cnt = (B - 1) * (2 * N - B - 1);
    if (cnt > 0) {
      k_cross<T, VEC, PK><<<cnt, 256, 0, st>>>(d, n, c, kb, B, N, prio);
      FWX_LAUNCHED("k_cross");
    }
  }
  return FWX_OK;
}

// Shape of the 128-wide rest kernel -- the register micro-tile: 2 = 512
// threads, 8x4 cells each, 126 VGPRs, two blocks (4 waves/SIMD) per CU; 0 =
// 256 threads, 8x8 cells, 32-deep chunks (2 waves/SIMD); 1 = 256 threads with
// 16-deep chunks (3 waves/SIMD, twice the barriers).  The tuned table picks
// it per size class (fwx_tune.cpp); FWX_REST_VARIANT=0/1/2 overrides it.
static int rest_variant(int32_t n) {
  const char* e = std::getenv("FWX_REST_VARIANT");
  return (e && e[0] >= '0' && e[0] <= '2') ? e[0] - '0' : tuned_rest_variant(n);
}

// Rest tiles outside the R-tile band [s0/R, s0/R + excl): excl covers stage
// s alone (plain schedule) or stage s and the next stage (look-ahead).
template <int R, bool VEC, bool PK>
static int rest_pass(int32_t* d, int32_t n, int32_t s0, int32_t S, int32_t excl, hipStream_t st) {
  if (t_gate) return ungated_launch("k_rest");
  const int nr = (n + R - 1) / R;
  const int sr0 = s0 / R;
  const int srn = min(excl, nr - sr0);
  const int m = nr - srn;
  if (m <= 0) return FWX_OK;
  const int64_t total = (int64_t)m * m;
  const int per = (int)((total + 7) / 8);
  const int v = R == 128 ? rest_variant(n) : 0;
  if (v == 1)
    k_rest<R, VEC, PK, 16><<<8 * per, 256, 0, st>>>(d, n, s0, S, sr0, srn, m, per);
  else if (v == 2)
    k_rest512<R, VEC, PK><<<8 * per, 512, 0, st>>>(d, n, s0, S, sr0, srn, m, per);
  else
    k_rest<R, VEC, PK><<<8 * per, 256, 0, st>>>(d, n, s0, S, sr0, srn, m, per);
  FWX_LAUNCHED("k_rest");
  return FWX_OK;
}

// FWX_BAND16=0 keeps the band pass in pk64 when the rest passes run pk16.
static bool band16_enabled() {
  const char* e = std::getenv("FWX_BAND16");
  return !(e && e[0] == '0');
}

template <int R, bool VEC, bool PK, bool P16 = false>
static int band_pass(int32_t* d, int32_t n, int32_t s0, int32_t S, int32_t en, int32_t fn_tiles,
                     hipStream_t st, int band_prio = 0) {
  const int nr = (n + R - 1) / R;
  const int e0 = s0 / R;
  const int f0 = e0 + en;
  const int fn = min(fn_tiles, nr - f0);
  const int m1 = nr - en;
  if (fn <= 0 || m1 <= 0) return FWX_OK;
  const int cnt = fn * m1 + (m1 - fn) * fn;
  if (cnt <= 0) return FWX_OK;
  if (P16 && R == 64 && S % 16 == 0 && band16_enabled() && band_wide() && e0 % 2 == 0 && en % 2 == 0 &&
      f0 % 2 == 0 && fn % 2 == 0 && m1 % 2 == 0)
    k_rest_band16w<VEC><<<cnt / 2, 256, 0, st>>>(d, n, s0, S, e0, en, f0, fn, m1, band_prio, t_gate);
  else if (P16 && R == 64 && S % 16 == 0 && band16_enabled())
    k_rest_band16<64, VEC><<<cnt, 256, 0, st>>>(d, n, s0, S, e0, en, f0, fn, m1, t_gate);
  else if (t_gate)
    return ungated_launch("k_rest_band");
  else
    k_rest_band<R, VEC, PK><<<cnt, 256, 0, st>>>(d, n, s0, S, e0, en, f0, fn, m1);
  FWX_LAUNCHED("k_rest_band");
  return FWX_OK;
}

// Chunk depth of the pk16 rest kernel: the tuned one for size n, or
// FWX_P16_KC=16 / 32.  16-deep chunks hold it to 116 VGPRs and 24 KB of LDS
// (four blocks per CU), 32-deep ones take 157 VGPRs and 52 KB (three).
static int p16_chunk(int32_t n) {
  const char* e = std::getenv("FWX_P16_KC");
  const int v = e ? std::atoi(e) : 0;
  return (v == 16 || v == 32) ? v : tuned_p16_chunk(n);
}

template <int R, bool VEC>
static int rest_pass16(int32_t* d, int32_t n, int32_t s0, int32_t S, int32_t excl, hipStream_t st) {
  const int nr = (n + R - 1) / R;
  const int sr0 = s0 / R;
  const int srn = min(excl, nr - sr0);
  const int m = nr - srn;
  if (m <= 0) return FWX_OK;
  const int64_t total = (int64_t)m * m;
  const int per = (int)((total + 7) / 8);
  if (t_pack && R == 128) {
    // this stage's packed panels (pack_stage ran after its lead rounds)
    const int b = (s0 / (t_pack->tile * t_pack->kappa)) % 3;
    // the last stage's rest pass leaves every cell outside the stage final
    int* capped = t_capped && s0 + S >= n ? t_capped : nullptr;
    if (rest16p_waves() == 5)
      k_rest16p<128, 5><<<8 * per, 256, 0, st>>>(d, n, t_pack->wa[b], t_pack->wb[b], S, sr0, srn, m, per, t_gate,
                                                 capped);
    else
      k_rest16p<128, 4><<<8 * per, 256, 0, st>>>(d, n, t_pack->wa[b], t_pack->wb[b], S, sr0, srn, m, per, t_gate,
                                                 capped);
    FWX_LAUNCHED("k_rest16p");
    return FWX_OK;
  }
  if (t_gate) return ungated_launch("k_rest16");
  if (p16_chunk(n) == 16)
    k_rest16<R, VEC, 16><<<8 * per, 256, 0, st>>>(d, n, s0, S, sr0, srn, m, per);
  else
    k_rest16<R, VEC, 32><<<8 * per, 256, 0, st>>>(d, n, s0, S, sr0, srn, m, per);
  FWX_LAUNCHED("k_rest16");
  return FWX_OK;
}

// P16: the 128-wide rest passes run in the pk16 form (the caller checks the
// result for saturated cells afterwards).
template <bool VEC, bool PK, bool P16>
static int rest_dispatch(int R, int32_t* d, int32_t n, int32_t s0, int32_t S, int32_t excl, hipStream_t st) {
  if (R == 128) return P16 ? rest_pass16<128, VEC>(d, n, s0, S, excl, st) : rest_pass<128, VEC, PK>(d, n, s0, S, excl, st);
  if (R == 64) return rest_pass<64, VEC, PK>(d, n, s0, S, excl, st);
  return rest_pass<32, VEC, PK>(d, n, s0, S, excl, st);
}

// The band is only ~2(kappa*T/n) of a stage's rest work, too few 128-wide
// tiles to fill the chip (about one per CU at n=8192), so it runs on 64-wide
// tiles (4x the blocks) unless FWX_BAND=128.
static bool band_narrow() {
  const char* e = std::getenv("FWX_BAND");
  return !(e && e[0] == '1');
}

template <bool VEC, bool PK, bool P16 = false>
static int band_dispatch(int R, int32_t* d, int32_t n, int32_t s0, int32_t S, int32_t en, int32_t fn,
                         hipStream_t st, int band_prio = 0) {
  if (R == 128) {
    if (band_narrow()) return band_pass<64, VEC, PK, P16>(d, n, s0, S, 2 * en, 2 * fn, st, band_prio);
    return band_pass<128, VEC, PK>(d, n, s0, S, en, fn, st);
  }
  if (R == 64) return band_pass<64, VEC, PK>(d, n, s0, S, en, fn, st);
  return band_pass<32, VEC, PK>(d, n, s0, S, en, fn, st);
}

// FWX_LEAD_BLOCK=0 keeps the kappa lead rounds in packed pk16 solves.
static bool lead_block_enabled() {
  const char* e = std::getenv("FWX_LEAD_BLOCK");
  return !(e && e[0] == '0');
}

// The block lead phase of one stage (k_block_square16, k_block_panel16):
// ceil(log2(B T)) squaring launches, of which those after convergence
// return at once, and one panel launch.  Replaces B lead rounds of (closure,
// panels, cross): the same panel work as products of depth B T instead of
// B products of depth T, in 1 + log2(B T) launches instead of 1 + 2 B.
// Depth parts per tile of the first stage's block squarings
// (k_block_square16s): FWX_SQ_SPLIT = 4 (default) or 8; 0 or 1 keeps one
// workgroup per tile.  The first stage's lead phase runs with nothing beside
// it, on B^2 = 64 workgroups at (64, 8).  Measured at n=8192 on (64, 8):
// 9.045-9.077 ms per step unsplit, 9.012-9.014 with 4 parts, 9.029-9.032
// with 8; splitting every stage's squarings (which run beside the rest
// passes) 9.00 against 8.94-8.97 for the first stage alone.
static int sq_split() {
  const char* e = std::getenv("FWX_SQ_SPLIT");
  const int v = e ? std::atoi(e) : 4;
  return (v == 4 || v == 8) ? v : 0;
}

template <bool VEC>
static int lead_block(int32_t* d, int32_t n, int32_t kb, int32_t B, int32_t N, int prio, hipStream_t st) {
  PackWs* w = t_pack;
  const uint32_t gen = ++w->gen;
  int nsq = 0;
  while ((1 << nsq) < B * 64) ++nsq;
  if (nsq > kMaxSquarings) return fail(FWX_ERR_INVALID_ARG, "stage block too wide for the block lead phase");
  // parts of depth B 64 / ks, a multiple of the 16-deep chunk
  const int ks = kb == 0 && n % 64 == 0 ? sq_split() : 0;
  const bool split = ks > 0 && (B * 64) % (16 * ks) == 0;
  for (int it = 0; it < nsq; ++it) {
    if (split && ks == 8)
      k_block_square16s<8, VEC><<<B * B * 8, 256, 0, st>>>(d, n, kb, B, it, w->flags, gen, prio, t_gate);
    else if (split)
      k_block_square16s<4, VEC><<<B * B * 4, 256, 0, st>>>(d, n, kb, B, it, w->flags, gen, prio, t_gate);
    else
      k_block_square16<64, VEC><<<B * B, 256, 0, st>>>(d, n, kb, B, it, w->flags, gen, prio, t_gate);
    FWX_LAUNCHED("k_block_square16");
  }
  if (N > B) {
    if (panel_wide() && B % 2 == 0 && kb % 2 == 0 && (N - B) % 2 == 0) {
      k_block_panel16w<VEC><<<B * (N - B), 256, 0, st>>>(d, n, kb, B, N, prio, t_gate);
      FWX_LAUNCHED("k_block_panel16w");
    } else {
      k_block_panel16<64, VEC><<<2 * B * (N - B), 256, 0, st>>>(d, n, kb, B, N, prio, t_gate);
      FWX_LAUNCHED("k_block_panel16");
    }
  }
  return FWX_OK;
}

// The lead rounds of one stage.
template <int T, bool VEC, bool PK, bool P16 = false>
static int lead_rounds(int32_t* d, int32_t n, int32_t kb, int32_t B, int32_t N, int prio, hipStream_t st) {
  if (P16 && T == 64 && t_pack && B > 1 && lead_block_enabled()) {
    if (int rc = lead_block<VEC>(d, n, kb, B, N, prio, st)) return rc;
  } else {
    for (int l = 0; l < B; ++l)
      if (int rc = lead_round<T, VEC, PK, P16>(d, n, kb + l, kb, B, N, prio, st)) return rc;
  }
  // the stage's panels are final: pack them for its rest pass
  if (P16 && t_pack) return pack_stage(d, n, kb, B * T, prio, st);
  return FWX_OK;
}

static double lead_work(int32_t n, int32_t T, int32_t kb, int32_t B) {
  const double L = (double)min(B * T, n - kb * T), nn = (double)n;
  return (nn * nn - (nn - L) * (nn - L)) * L;
}

// The lead rounds of one stage, bracketed for the profiler.
template <int T, bool VEC, bool PK, bool P16 = false>
static int lead_stage(int32_t* d, int32_t n, int32_t kb, int32_t B, int32_t N, int prio, hipStream_t st) {
  hipEvent_t ev;
  if (int rc = prof_open(st, &ev)) return rc;
  if (int rc = lead_rounds<T, VEC, PK, P16>(d, n, kb, B, N, prio, st)) return rc;
  return prof_close(st, ev, FWX_PROF_LEAD, lead_work(n, T, kb, B));
}

// Plain schedule: lead(s), rest(s), lead(s+1), ... on one stream.  While
// profiling, one event between adjacent groups ends one and starts the next.
template <int T, bool VEC, bool PK, bool P16>
static int run_blocked(int32_t* d, int32_t n, int32_t kappa, hipStream_t st) {
  const int N = (n + T - 1) / T;
  const int span = kappa * T;
  const int R = (span % 128 == 0) ? 128 : (span % 64 == 0) ? 64 : 32;
  hipEvent_t ev;
  if (int rc = prof_open(st, &ev)) return rc;
  for (int kb = 0; kb < N; kb += kappa) {
    const int B = min(kappa, N - kb);
    const int s0 = kb * T, S = B * T;
    const double L = (double)min(S, n - s0), nn = (double)n;
    if (int rc = lead_rounds<T, VEC, PK, P16>(d, n, kb, B, N, 0, st)) return rc;
    if (int rc = prof_chain(st, &ev, FWX_PROF_LEAD, lead_work(n, T, kb, B))) return rc;
    if (int rc = rest_dispatch<VEC, PK, P16>(R, d, n, s0, S, (S + R - 1) / R, st)) return rc;
    if (int rc = prof_chain(st, &ev, FWX_PROF_REST, (nn - L) * (nn - L) * L)) return rc;
  }
  return FWX_OK;
}

// FWX_LEAD_PRIO=0 keeps the look-ahead's lead kernels at the default wave
// priority (read on every call).
static bool lead_priority_enabled() {
  const char* e = std::getenv("FWX_LEAD_PRIO");
  return !(e && e[0] == '0');
}

// Look-ahead schedule.  Stage s+1's lead rounds only read and write tiles
// with a row or column in band s+1; stage s's rest pass is split into the
// tiles of that cross (rest_band, first, on the caller's stream) and the
// others (after it).  A side stream runs lead(s+1) as soon as rest_band(s)
// is done, concurrently with the other rest tiles of stage s, which touch
// neither band s+1 nor anything lead(s+1) writes (lead(s+1) writes only band
// s+1 rows/columns; rest(s) reads only band s rows/columns).  The caller's
// stream waits for lead(s+1) before stage s+1's rest.
template <int T, bool VEC, bool PK, bool P16>
static int run_blocked_lookahead(int32_t* d, int32_t n, int32_t kappa, hipStream_t st, hipStream_t side,
                                 hipEvent_t evA, hipEvent_t evB) {
  const int N = (n + T - 1) / T;
  const int span = kappa * T;
  const int R = (span % 128 == 0) ? 128 : (span % 64 == 0) ? 64 : 32;
  const double nn = (double)n;
  const int prio = lead_priority_enabled() ? 1 : 0;
  // While profiling, `ev` chains the caller's stream: the event that ends
  // one group starts the next (so a band group also holds the wait for the
  // side stream's lead), and the band's end event is the one the side stream
  // waits on.  Two markers per stage on the caller's stream, as unprofiled.
  hipEvent_t ev;
  if (int rc = prof_open(st, &ev)) return rc;
  const int B0 = min(kappa, N);
  if (int rc = lead_rounds<T, VEC, PK, P16>(d, n, 0, B0, N, 0, st)) return rc;
  if (int rc = prof_chain(st, &ev, FWX_PROF_LEAD, lead_work(n, T, 0, B0))) return rc;
  for (int kb = 0; kb < N; kb += kappa) {
    const int B = min(kappa, N - kb);
    const int s0 = kb * T, S = B * T;
    const double L = (double)min(S, n - s0);
    const int en = (S + R - 1) / R;
    const int kb2 = kb + kappa;
    if (kb2 < N) {
      const int B2 = min(kappa, N - kb2);
      const int fn = (B2 * T + R - 1) / R;
      const double L2 = (double)min(B2 * T, n - kb2 * T);
      if (int rc = band_dispatch<VEC, PK, P16>(R, d, n, s0, S, en, fn, st)) return rc;
      if (int rc = prof_chain(st, &ev, FWX_PROF_BAND, ((nn - L) * (nn - L) - (nn - L - L2) * (nn - L - L2)) * L))
        return rc;
      if (!ev) FWX_HIP(hipEventRecord(evA, st));
      FWX_HIP(hipStreamWaitEvent(side, ev ? ev : evA, 0));
      if (int rc = lead_stage<T, VEC, PK, P16>(d, n, kb2, B2, N, prio, side)) return rc;
      FWX_HIP(hipEventRecord(evB, side));
      if (int rc = rest_dispatch<VEC, PK, P16>(R, d, n, s0, S, en + fn, st)) return rc;
      if (int rc = prof_chain(st, &ev, FWX_PROF_REST, (nn - L - L2) * (nn - L - L2) * L)) return rc;
      FWX_HIP(hipStreamWaitEvent(st, evB, 0));
    } else {
      if (int rc = rest_dispatch<VEC, PK, P16>(R, d, n, s0, S, en, st)) return rc;
      if (int rc = prof_chain(st, &ev, FWX_PROF_REST, (nn - L) * (nn - L) * L)) return rc;
    }
  }
  return FWX_OK;
}

// Side-band schedule: the caller's stream runs only rest passes.  The side
// stream runs, per stage s, band(s) -- the rest(s) tiles with a row or a
// column in band s+1 -- and then lead(s+1); the caller's stream runs the
// other rest(s) tiles.  The two write disjoint cells and neither writes what
// the other reads: rest_other(s) reads band s rows and columns outside band
// s+1 and writes cells outside bands s and s+1; band(s) and lead(s+1) write
// only cells with a row or a column in band s+1 (none in band s), and read
// band s and band s+1.  Ordering: rest_other(s) waits for lead(s) (event L);
// band(s+1) waits for rest_other(s) (event R), which wrote the cells of the
// band s+2 cross that band(s+1) updates next.
template <int T, bool VEC, bool PK, bool P16>
static int run_blocked_sideband(int32_t* d, int32_t n, int32_t kappa, hipStream_t st, hipStream_t side,
                                hipEvent_t evR, hipEvent_t evL) {
  const int N = (n + T - 1) / T;
  const int span = kappa * T;
  const int R = (span % 128 == 0) ? 128 : (span % 64 == 0) ? 64 : 32;
  const double nn = (double)n;
  const int prio = lead_priority_enabled() ? 1 : 0;
  FWX_HIP(hipEventRecord(evR, st));  // the side stream starts after the caller's prior work
  FWX_HIP(hipStreamWaitEvent(side, evR, 0));
  if (int rc = lead_stage<T, VEC, PK, P16>(d, n, 0, min(kappa, N), N, prio, side)) return rc;
  FWX_HIP(hipEventRecord(evL, side));
  hipEvent_t ev = nullptr;  // profiling markers of the rest passes on the caller's stream
  for (int kb = 0; kb < N; kb += kappa) {
    const int B = min(kappa, N - kb);
    const int s0 = kb * T, S = B * T;
    const double L = (double)min(S, n - s0);
    const int en = (S + R - 1) / R;
    const int kb2 = kb + kappa;
    FWX_HIP(hipStreamWaitEvent(st, evL, 0));
    // a fresh start marker after the wait, so that a profiled rest group
    // times the rest kernel alone, not the wait for the side stream
    if (int rc = prof_open(st, &ev)) return rc;
    if (kb2 < N) {
      const int B2 = min(kappa, N - kb2);
      const int fn = (B2 * T + R - 1) / R;
      const double L2 = (double)min(B2 * T, n - kb2 * T);
      hipEvent_t es;
      if (int rc = prof_open(side, &es)) return rc;
      if (int rc = band_dispatch<VEC, PK, P16>(R, d, n, s0, S, en, fn, side)) return rc;
      if (int rc = prof_chain(side, &es, FWX_PROF_BAND, ((nn - L) * (nn - L) - (nn - L - L2) * (nn - L - L2)) * L))
        return rc;
      if (int rc = lead_rounds<T, VEC, PK, P16>(d, n, kb2, B2, N, prio, side)) return rc;
      if (int rc = prof_close(side, es, FWX_PROF_LEAD, lead_work(n, T, kb2, B2))) return rc;
      FWX_HIP(hipEventRecord(evL, side));
      if (int rc = rest_dispatch<VEC, PK, P16>(R, d, n, s0, S, en + fn, st)) return rc;
      if (int rc = prof_chain(st, &ev, FWX_PROF_REST, (nn - L - L2) * (nn - L - L2) * L)) return rc;
      if (!ev) FWX_HIP(hipEventRecord(evR, st));
      FWX_HIP(hipStreamWaitEvent(side, ev ? ev : evR, 0));
    } else {
      if (int rc = rest_dispatch<VEC, PK, P16>(R, d, n, s0, S, en, st)) return rc;
      if (int rc = prof_chain(st, &ev, FWX_PROF_REST, (nn - L) * (nn - L) * L)) return rc;
    }
  }
  return FWX_OK;
}

// Raised wave priority for the 64 x 128 pk16 band kernel in the
// side-band-ahead schedule (FWX_BAND_PRIO, default 1): bit 0 for the X(s)
// passes on the caller's stream, bit 1 for the band passes on the side
// stream.  X(s) shares the SIMDs with the side stream's lead kernels, which
// raise theirs; at the same priority it finishes sooner and the caller's
// stream -- the longer of the two -- waits less.  Measured at n=8192 on
// (64, 8): 9.26 ms per step at 0, 9.04 at 1, 9.95 at 2, 9.82 at 3.
static int band_prio_mode() {
  const char* e = std::getenv("FWX_BAND_PRIO");
  return e ? std::atoi(e) & 3 : 1;
}

// Side-band-ahead schedule (FWX_LOOKAHEAD=3): the side-band schedule with
// the side stream one stage ahead.  Measured at n=8192 on (64, 4): 17.31-17.34
// vs 17.39-17.40 ms per step for the side-band schedule -- the step is
// already within ~1 % of n^3 at the rest kernel's own rate, so removing the
// caller's waits for the side stream gains little.  Stage s's rest pass outside bands s and s+1 is split once
// more: X(s), its tiles with a row or a column in band s+2, runs first on the
// caller's stream (on 64-wide tiles, like the band), then Y(s), every other
// tile.  band(s+1) -- the rest(s+1) tiles with a row or a column in band s+2
// -- reads and writes only cells that rest(s) wrote in X(s), so it waits for
// X(s) (event X) instead of the whole rest pass; the side stream then runs
// band(s+1) and lead(s+2) while the caller's stream still runs Y(s), and the
// event the caller waits for before stage s+2 (L) has completed by then.
// Read/write sets: Y(s) writes cells outside bands s, s+1, s+2 and reads
// band-s panel cells outside band s+2; band(s+1) writes cells with a row or
// column in band s+2 outside band s+1 and reads band s+1; lead(s+2) reads
// and writes band s+2 only; X(s) writes cells with a row or column in band
// s+2 outside bands s and s+1, which the side stream touches only after
// event X.  Ordering: X(s), Y(s) wait for lead(s) (event L); band(s+1) waits
// for X(s) (event X).  Each stage enqueues X(s) before its side-stream work:
// the host takes a few microseconds per launch, and with the side stream's
// band and lead launches first, X(0) started only after lead(1) had run
// (~160 us per solve at n=8192).
template <int T, bool VEC, bool PK, bool P16>
static int run_blocked_ahead(int32_t* d, int32_t n, int32_t kappa, hipStream_t st, hipStream_t side,
                             const hipEvent_t (&evX)[2], const hipEvent_t (&evL)[2]) {
  const int N = (n + T - 1) / T;
  const int span = kappa * T;
  const int R = (span % 128 == 0) ? 128 : (span % 64 == 0) ? 64 : 32;
  const double nn = (double)n;
  const int prio = lead_priority_enabled() ? 1 : 0;
  // Stage s's handoffs use evL[s & 1] (lead(s) done) and evX[s & 1] (X(s)
  // done): each is recorded once per two stages, and every wait on it is
  // enqueued before its next record.
  FWX_HIP(hipEventRecord(evX[1], st));  // the side stream starts after the caller's prior work
  FWX_HIP(hipStreamWaitEvent(side, evX[1], 0));
  if (int rc = lead_stage<T, VEC, PK, P16>(d, n, 0, min(kappa, N), N, prio, side)) return rc;
  FWX_HIP(hipEventRecord(evL[0], side));
  hipEvent_t ev = nullptr;  // profiling markers of the caller's stream
  int si = 0;
  for (int kb = 0; kb < N; kb += kappa, ++si) {
    const int B = min(kappa, N - kb);
    const int s0 = kb * T, S = B * T;
    const double L = (double)min(S, n - s0);
    const int en = (S + R - 1) / R;  // band s in R tiles
    const int kb2 = kb + kappa, kb3 = kb2 + kappa;
    FWX_HIP(hipStreamWaitEvent(st, evL[si & 1], 0));
    if (int rc = prof_open(st, &ev)) return rc;
    if (kb2 < N) {
      const int B2 = min(kappa, N - kb2);
      const int fn = (B2 * T + R - 1) / R;
      const double L2 = (double)min(B2 * T, n - kb2 * T);
      const double outside2 = nn - L - L2;  // rows outside bands s and s+1
      int gn = 0;
      double L3 = 0;
      if (kb3 < N) {
        // caller: X(s), the rest tiles with a row or column in band s+2,
        // enqueued before this stage's side-stream work so that the host
        // does not hold it back behind those launches
        const int B3 = min(kappa, N - kb3);
        gn = (B3 * T + R - 1) / R;
        L3 = (double)min(B3 * T, n - kb3 * T);
        if (int rc = band_dispatch<VEC, PK, P16>(R, d, n, s0, S, en + fn, gn, st, band_prio_mode() & 1)) return rc;
        if (int rc = prof_chain(st, &ev, FWX_PROF_BAND, (outside2 * outside2 - (outside2 - L3) * (outside2 - L3)) * L))
          return rc;
        FWX_HIP(hipEventRecord(evX[si & 1], st));
      }
      // side: band(s) -- after X(s-1), which wrote the band s+1 cells it
      // updates -- and lead(s+1)
      if (si > 0) FWX_HIP(hipStreamWaitEvent(side, evX[(si - 1) & 1], 0));
      hipEvent_t es;
      if (int rc = prof_open(side, &es)) return rc;
      if (int rc = band_dispatch<VEC, PK, P16>(R, d, n, s0, S, en, fn, side, (band_prio_mode() >> 1) & 1)) return rc;
      if (int rc = prof_chain(side, &es, FWX_PROF_BAND, ((nn - L) * (nn - L) - (nn - L - L2) * (nn - L - L2)) * L))
        return rc;
      if (int rc = lead_rounds<T, VEC, PK, P16>(d, n, kb2, B2, N, prio, side)) return rc;
      if (int rc = prof_close(side, es, FWX_PROF_LEAD, lead_work(n, T, kb2, B2))) return rc;
      FWX_HIP(hipEventRecord(evL[(si + 1) & 1], side));
      // caller: Y(s), every other rest tile
      if (gn > 0) {
        if (int rc = rest_dispatch<VEC, PK, P16>(R, d, n, s0, S, en + fn + gn, st)) return rc;
        if (int rc = prof_chain(st, &ev, FWX_PROF_REST, (outside2 - L3) * (outside2 - L3) * L)) return rc;
      } else {
        if (int rc = rest_dispatch<VEC, PK, P16>(R, d, n, s0, S, en + fn, st)) return rc;
        if (int rc = prof_chain(st, &ev, FWX_PROF_REST, outside2 * outside2 * L)) return rc;
      }
    } else {
      if (int rc = rest_dispatch<VEC, PK, P16>(R, d, n, s0, S, en, st)) return rc;
      if (int rc = prof_chain(st, &ev, FWX_PROF_REST, (nn - L) * (nn - L) * L)) return rc;
    }
  }
  return FWX_OK;
}

// One non-blocking side stream per device (default priority), created on
// first use and kept for the life of the process.
// While this thread captures a solve into a graph, its side stream is a
// private one (t_capture_side), so that no other thread's work on the shared
// side stream can enter, or be refused by, the capture.
static thread_local hipStream_t t_capture_side = nullptr;

int side_stream(hipStream_t* out) {
  if (t_capture_side) {
    *out = t_capture_side;
    return FWX_OK;
  }
  static std::mutex mu;
  static hipStream_t streams[64] = {};
  int dev = 0;
  FWX_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) return fail(FWX_ERR_UNSUPPORTED, "device index above 63");
  std::lock_guard<std::mutex> lk(mu);
  if (!streams[dev]) FWX_HIP(hipStreamCreateWithFlags(&streams[dev], hipStreamNonBlocking));
  *out = streams[dev];
  return FWX_OK;
}

// The schedule: the tuned one of the size class (2 or 3), or FWX_LOOKAHEAD
// (read on every call): 0 = single stream, 1 = look-ahead with the band on
// the caller's stream, 2 = side-band, 3 = side-band with the side stream a
// stage ahead.
static int lookahead_mode(int32_t n) {
  const char* e = std::getenv("FWX_LOOKAHEAD");
  return (e && e[0] >= '0' && e[0] <= '3') ? e[0] - '0' : tuned_lookahead(n);
}

template <int T, bool VEC, bool PK, bool P16>
static int run_any(int32_t* d, int32_t n, int32_t kappa, hipStream_t st) {
  const int N = (n + T - 1) / T;
  const int mode = lookahead_mode(n);
  if (mode == 0 || N <= kappa) return run_blocked<T, VEC, PK, P16>(d, n, kappa, st);
  hipStream_t side;
  if (int rc = side_stream(&side)) return rc;
  // ordering events of this call (evs[2], evs[3] for the side-band-ahead
  // schedule's two-stage rings)
  hipEvent_t evs[4] = {};
  const int nev = mode == 3 ? 4 : 2;
  for (int i = 0; i < nev; ++i) {
    const hipError_t e = hipEventCreateWithFlags(&evs[i], event_flags(false));
    if (e != hipSuccess) {
      for (int k = 0; k < i; ++k) (void)hipEventDestroy(evs[k]);
      return hip_fail(e, "hipEventCreateWithFlags", __FILE__, __LINE__);
    }
  }
  const hipEvent_t evX[2] = {evs[0], evs[1]}, evL[2] = {evs[2], evs[3]};
  const int rc = mode == 3   ? run_blocked_ahead<T, VEC, PK, P16>(d, n, kappa, st, side, evX, evL)
                 : mode == 2 ? run_blocked_sideband<T, VEC, PK, P16>(d, n, kappa, st, side, evs[0], evs[1])
                             : run_blocked_lookahead<T, VEC, PK, P16>(d, n, kappa, st, side, evs[0], evs[1]);
  for (int i = 0; i < nev; ++i) (void)hipEventDestroy(evs[i]);
  return rc;
}

template <bool VEC, bool PK, bool P16>
static int by_tile(int32_t* d, int32_t n, int32_t tile, int32_t kappa, hipStream_t st) {
  switch (tile) {
    case 32:
      return run_any<32, VEC, PK, P16>(d, n, kappa, st);
    case 64:
      return run_any<64, VEC, PK, P16>(d, n, kappa, st);
    case 128:
      return run_any<128, VEC, PK, P16>(d, n, kappa, st);
    default:
      return fail(FWX_ERR_INVALID_ARG, "tile must be 32, 64 or 128");
  }
}

// form: 0 = min3, 1 = pk64, 2 = pk64 with the 128-wide rest passes in pk16.
static int enqueue_direct(int32_t* d, int32_t n, int32_t tile, int32_t kappa, bool vec, int form,
                          hipStream_t st) {
  if (vec) {
    if (form == 2) return by_tile<true, true, true>(d, n, tile, kappa, st);
    return form ? by_tile<true, true, false>(d, n, tile, kappa, st) : by_tile<true, false, false>(d, n, tile, kappa, st);
  }
  if (form == 2) return by_tile<false, true, true>(d, n, tile, kappa, st);
  return form ? by_tile<false, true, false>(d, n, tile, kappa, st) : by_tile<false, false, false>(d, n, tile, kappa, st);
}

// ---------------------------------------------------------------------------
// Small n: tile 32 with kappa 1 runs as one fused launch per k-tile round
// (fwx_small.hip), each block picking the relaxation form from the operands
// it loaded, so there is no range pass and the host never waits.  Up to
// kSmallMaxN; FWX_SMALL=0 keeps the three launches per round.
constexpr int32_t kSmallMaxN = 4096;

static bool small_path(int32_t n, int32_t tile, int32_t kappa) {
  if (tile != 32 || kappa != 1 || n > kSmallMaxN) return false;
  const char* e = std::getenv("FWX_SMALL");
  return !(e && e[0] == '0');
}

// The rounds, bracketed for the profiler as one lead group of n^3
// relaxations.
static int enqueue_small(int32_t* d, int32_t n, hipStream_t st) {
  hipEvent_t ev;
  if (int rc = prof_open(st, &ev)) return rc;
  if (int rc = enqueue_small_rounds(d, n, st)) return rc;
  const double nn = (double)n;
  return prof_close(st, ev, FWX_PROF_LEAD, nn * nn * nn);
}

// ---------------------------------------------------------------------------
// Replay cache.  A solve is a fixed sequence of a few hundred launches for a
// given (buffer, n, tile, kappa, form, stream); the first call captures it
// (both streams of the look-ahead schedule, joined by events) into a
// hipGraph and later calls replay it, which removes most inter-launch gaps.
// Entries are kept per process, at most kGraphCache, least recently used
// first out.  FWX_GRAPH=0 disables it.  Used up to n = kGraphMaxN only:
// measured on MI355X (tuned tile/kappa, best of 4-10 solves) the replay is
// 10 % faster than direct launches at n=2048 and 2 % at n=4096, but 1.7 %
// slower at n=8192, where launch gaps are negligible next to ~0.5 ms rest
// kernels.  The small-n rounds launch directly (FWX_SMALL_GRAPH=1 replays
// them): one launch per round, and with the look-ahead closure by squaring
// the direct launches measured 1-20 % faster than the replay at every n
// from 64 to 1024 and equal at 1536-2048.  Profiled calls launch directly:
// events recorded by a graph's event nodes cannot be timed with
// hipEventElapsedTime (the runtime reports an invalid handle).

namespace {

struct GraphKey {
  int32_t* d;
  int32_t n, tile, kappa;
  bool vec;
  int form;  // 0 min3, 1 pk64, 2 pk64 with pk16 rest passes
  bool small;  // the fused small-n rounds (each block then picks the form itself)
  int small_cfg;  // their region width and threads (small_region_knob)
  int la;
  int knobs;  // FWX_DPANEL / FWX_REST_VARIANT / FWX_BAND / FWX_DIAG_NTH / FWX_LEAD_PRIO choices
  hipStream_t st;
  int dev;
  bool operator==(const GraphKey& o) const {
    return d == o.d && n == o.n && tile == o.tile && kappa == o.kappa && vec == o.vec && form == o.form &&
           small == o.small && small_cfg == o.small_cfg && la == o.la && knobs == o.knobs && st == o.st &&
           dev == o.dev;
  }
};

struct GraphEntry {
  GraphKey key;
  hipGraph_t graph;
  hipGraphExec_t exec;
  uint64_t used;
};

constexpr size_t kGraphCache = 8;
constexpr int32_t kGraphMaxN = 4096;
std::mutex g_graph_mu;
std::vector<GraphEntry> g_graphs;
uint64_t g_graph_clock = 0;

bool graphs_enabled() {
  const char* e = std::getenv("FWX_GRAPH");
  return !(e && e[0] == '0');
}

bool small_graphs_enabled() {
  const char* e = std::getenv("FWX_SMALL_GRAPH");
  return e && e[0] == '1';
}

}  // namespace

// The pk16 rest passes (fwx_tile16.hpp) run whenever the pk64 form is exact
// and the rest pass is 128 wide; FWX_PK16=0 keeps the pk64 form throughout.
static bool pk16_enabled() {
  const char* e = std::getenv("FWX_PK16");
  return !(e && e[0] == '0');
}

static int rest_width(int32_t tile, int32_t kappa) {
  const int span = kappa * tile;
  return (span % 128 == 0) ? 128 : (span % 64 == 0) ? 64 : 32;
}

// One solve in the given form: launched directly, or replayed from the graph
// cache (captured on first use).
static int run_form(int32_t* d, int32_t n, int32_t tile, int32_t kappa, bool vec, int form, bool small,
                    bool direct, hipStream_t st) {
  if (direct) return small ? enqueue_small(d, n, st) : enqueue_direct(d, n, tile, kappa, vec, form, st);
  int dev = 0;
  FWX_HIP(hipGetDevice(&dev));
  const int knobs = (fused_dpanel() ? 1 : 0) | (rest_variant(n) << 1) | (band_narrow() ? 8 : 0) | (diag_threads() << 4) |
                    (lead_priority_enabled() ? 1 << 20 : 0) | (p16_chunk(n) == 16 ? 1 << 21 : 0) |
                    (diag_sq_enabled() ? 1 << 22 : 0) | (lead_close_enabled() ? 1 << 23 : 0) |
                    (band_wide() ? 1 << 24 : 0) | (panel_wide() ? 1 << 25 : 0) | (band_prio_mode() << 26);
  const GraphKey key{d, n, tile, kappa, vec, form, small, small ? small_region_knob(n) : 0, lookahead_mode(n), knobs, st,
                     dev};
  std::lock_guard<std::mutex> lk(g_graph_mu);
  auto replay = [&](GraphEntry& e) -> int {
    e.used = ++g_graph_clock;
    FWX_HIP(hipGraphLaunch(e.exec, st));
    return FWX_OK;
  };
  for (GraphEntry& e : g_graphs)
    if (e.key == key) return replay(e);
  // capture on a private stream (`st` may be the legacy default stream),
  // with a private side stream for the two-stream schedules
  hipStream_t cap, cap_side;
  FWX_HIP(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
  const hipError_t es = hipStreamCreateWithFlags(&cap_side, hipStreamNonBlocking);
  if (es != hipSuccess) {
    (void)hipStreamDestroy(cap);
    return hip_fail(es, "hipStreamCreateWithFlags", __FILE__, __LINE__);
  }
  hipError_t e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
  int rc = FWX_OK;
  t_capture_side = cap_side;
  if (e == hipSuccess) rc = small ? enqueue_small(d, n, cap) : enqueue_direct(d, n, tile, kappa, vec, form, cap);
  t_capture_side = nullptr;
  hipGraph_t graph = nullptr;
  const hipError_t e2 = hipStreamEndCapture(cap, &graph);
  (void)hipStreamDestroy(cap);
  (void)hipStreamDestroy(cap_side);
  if (e != hipSuccess) return hip_fail(e, "hipStreamBeginCapture", __FILE__, __LINE__);
  if (rc || e2 != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc ? rc : hip_fail(e2, "hipStreamEndCapture", __FILE__, __LINE__);
  }
  hipGraphExec_t exec = nullptr;
  e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGraphDestroy(graph);
    return hip_fail(e, "hipGraphInstantiate", __FILE__, __LINE__);
  }
  if (g_graphs.size() >= kGraphCache) {
    size_t lru = 0;
    for (size_t i = 1; i < g_graphs.size(); ++i)
      if (g_graphs[i].used < g_graphs[lru].used) lru = i;
    (void)hipGraphExecDestroy(g_graphs[lru].exec);
    (void)hipGraphDestroy(g_graphs[lru].graph);
    g_graphs.erase(g_graphs.begin() + (long)lru);
  }
  g_graphs.push_back(GraphEntry{key, graph, exec, 0});
  return replay(g_graphs.back());
}

int enqueue_blocked_exact(int32_t* d, int32_t n, hipStream_t st) {
  int32_t tile = 0, kappa = 0;
  tuned_config(n, &tile, &kappa);
  const bool vec = (n % 4 == 0) && ((reinterpret_cast<uintptr_t>(d) & 15u) == 0);
  if (small_path(n, tile, kappa)) return enqueue_small(d, n, st);
  return enqueue_direct(d, n, tile, kappa, vec, 1, st);
}

static thread_local int t_last_form = -1;
static thread_local int t_last_exact_finish = 0;
static thread_local int t_last_packed = 0;

int last_solve_packed() { return t_last_packed; }

void last_solve_form(int32_t* form, int32_t* exact_finish) {
  *form = t_last_form;
  *exact_finish = t_last_exact_finish;
}

// FWX_RANGE_DEFER=0 makes every packed solve wait for its range check
// before enqueuing its kernels.
static bool range_defer_enabled() {
  const char* e = std::getenv("FWX_RANGE_DEFER");
  return !(e && e[0] == '0');
}

// A packed pk16 solve whose every launch goes to a gated kernel: 64-wide
// tiles with kappa > 1 and no one-tile last stage (every stage takes the
// block lead phase, never the lead rounds), the band in pk16 on 64-wide
// tiles, the rest passes on packed panels (128-wide, which the packed solve
// implies).  Such a solve enqueues its range check and its kernels without
// waiting (solve_deferred).
static bool range_deferrable(int32_t n, int32_t tile, int32_t kappa) {
  const int32_t N = n / 64;
  return range_defer_enabled() && tile == 64 && kappa > 1 && N % kappa != 1 && lead_block_enabled() &&
         band_narrow() && band16_enabled();
}

// FWX_CAP_FUSED=0 checks a deferred solve's result for cells at the pk16 cap
// by a pass over the matrix (k_at_least) instead of in its last stage's pack
// and rest pass.
static bool cap_fused() {
  const char* e = std::getenv("FWX_CAP_FUSED");
  return !(e && e[0] == '0');
}

// A packed pk16 solve with its range check deferred.  The check ORs into
// ws->gate[0] on st ahead of the solve's kernels, which return at entry when
// it is set; after the solve one read-back brings it with the cap check
// (ws->gate[1]: the last stage's k_pack16 reads every cell with a row or a
// column in that stage after its last write, and the last rest pass holds
// every other cell's final value).  Outside the range, d still holds the
// input and is solved in the min3 form; with a cell at the cap, the exact
// pk64 finish follows, as in the waiting path.  The workspace is held until
// the read-back.
static int solve_deferred(int32_t* d, int32_t n, int32_t tile, int32_t kappa, bool vec, PackWs* ws,
                          hipStream_t st) {
  const int64_t count = (int64_t)n * n;
  const bool fused = cap_fused();
  int rc = range_check_enqueue(d, count, ws->gate, st);
  if (!rc) {
    t_pack = ws;
    t_gate = ws->gate;
    t_capped = fused ? ws->gate + 1 : nullptr;
    rc = run_form(d, n, tile, kappa, vec, 2, false, true, st);
    t_pack = nullptr;
    t_gate = nullptr;
    t_capped = nullptr;
  }
  int h[2] = {1, 1};
  if (!rc) rc = deferred_checks_read(d, count, (int32_t)kP16Cap, !fused, ws->gate, st, h);
  if (rc) {
    (void)hipStreamSynchronize(st);
    pack_release(ws);
    return rc;
  }
  pack_release(ws);
  if (h[0] != 0) {
    t_last_form = 0;
    return run_form(d, n, tile, kappa, vec, 0, false, true, st);
  }
  t_last_form = 2;
  t_last_packed = 1;
  if (h[1] == 0) return FWX_OK;
  t_last_exact_finish = 1;
  return run_form(d, n, tile, kappa, vec, 1, false, true, st);
}

int enqueue_blocked(int32_t* d, int32_t n, int32_t tile, int32_t kappa, hipStream_t st) {
  const bool vec = (n % 4 == 0) && ((reinterpret_cast<uintptr_t>(d) & 15u) == 0);
  const bool small = small_path(n, tile, kappa);
  const bool direct = n > kGraphMaxN || !graphs_enabled() || profiling();
  t_last_form = -1;
  t_last_exact_finish = 0;
  t_last_packed = 0;
  if (small) return run_form(d, n, tile, kappa, vec, 0, true, direct || !small_graphs_enabled(), st);
  // Direct pk16 solves of whole 128-wide tiles with 16-deep chunks run their
  // rest passes on packed stage panels; the workspace is held until the
  // check after the solve has synchronised the stream.
  const bool packable = pk16_enabled() && rest_width(tile, kappa) == 128 && direct && vec && n % 128 == 0 &&
                        p16_chunk(n) == 16 && pack16_enabled();
  if (packable && range_deferrable(n, tile, kappa))
    if (PackWs* ws = pack_acquire(n, tile, kappa)) return solve_deferred(d, n, tile, kappa, vec, ws, st);
  bool pk = false;
  if (int rc = values_in_pk64_range(d, (int64_t)n * n, st, &pk)) return rc;
  const int form = !pk ? 0 : (pk16_enabled() && rest_width(tile, kappa) == 128) ? 2 : 1;
  t_last_form = form;
  PackWs* ws = nullptr;
  if (form == 2 && packable) ws = pack_acquire(n, tile, kappa);
  t_pack = ws;
  t_last_packed = ws ? 1 : 0;
  int rc = run_form(d, n, tile, kappa, vec, form, false, direct, st);
  t_pack = nullptr;
  if (rc) {
    if (ws) (void)hipStreamSynchronize(st);
    pack_release(ws);
    return rc;
  }
  if (form != 2) return FWX_OK;
  // Every cell below the pk16 cap is exact.  If any cell is at or above it,
  // the matrix (walk lengths of the input graph, each at most its input
  // value) is solved once more in the exact pk64 form: its closure equals
  // the input's.
  bool below = false;
  rc = values_below(d, (int64_t)n * n, (int32_t)kP16Cap, st, &below);
  if (rc && ws) (void)hipStreamSynchronize(st);
  pack_release(ws);
  if (rc) return rc;
  if (below) return FWX_OK;
  t_last_exact_finish = 1;
  return run_form(d, n, tile, kappa, vec, 1, false, direct, st);
}

}  // namespace fwx
